"""kinematics1d_oracle.py -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the reference's 1-D internal kinematics,
kinematic_modul (1d/kinematic_modul.c:1-247), used by the tests as the
checker of hw1d_kinematics; the product library never imports it.  It takes
the host setup of hawassi_c_amd.kinematics1d.Kinematics1D (levels, window,
sinhcosh2IP tables) and follows the reference's loops:

  H = eta + depth, gamma_-(H), gamma_+(H) by natural cubic splines of the H
  table (gsl_interp_cspline; here scipy's natural CubicSpline),
  cosh_kH / tanh_kH as their weighted sums at the table's end depths,
  phi^ = (Im u^ / k, -Re u^ / k), 0 at k = 0            (:97-113)
  gradient of H and depth (1d/functions.c:1850-1856)     (:115-116)
  per level z: cosh / sinh(k DZ), flat bottom (minD == maxD) or one DZ over
  the level: of k (z + depth[0]); else DZ = depth + z clipped at 0, weights
  from the DZ table with the exact ends (DZ == DZmin: (1, 0), DZ == DZplus:
  (0, 1)), at the level's own DZmin / DZplus              (:129-153)
  spectra factor, dtPhi, dzPhi, dxPhi and their real inverse transforms
  (ifft_complex_real_nx: Re of the backward DFT / n)      (:155-171)
  NaN above eta and below the bottom; dtPhi + deta factor,
  dxPhi + grad H factor + grad D dzPhi                    (:173-185)
  window x[indxIP1 ..], P = -dtPhi - g z - (dxPhi^2 + 2 dzPhi) / 2 (:186-201)
  accelerations (v - v_prev) / (t - tprev)                (:212-217)

As in the reference, node index j also indexes the wavenumber k[j].

How it is pinned (tests/test_kinematics1d.py): the linear-wave closed forms
on a flat bottom at rest (eta = 0 at a table node, so the H weights are
exact): dxPhi = -A k sin(kx) cosh(k(z+h))/cosh(kh), dzPhi = A k cos(kx)
sinh(k(z+h))/cosh(kh), dtPhi = B cos(kx) cosh(k(z+h))/cosh(kh) for
phi = A cos(kx), phi_t = B cos(kx), with the masks above the surface and
below the bottom; two planted misreadings (`plant`) fail them.  The
variable-depth weights of DZ are pinned through sinhcosh2IP's identity
(the weights reproduce cosh and sinh at the peak wavenumber) and the exact
ends; their use at the level's own DZmin / DZplus follows the reference's
text and is compared with the device only.

Planted mistakes (0 = none):
  1  phi^ = +i u^ / k (sign of the potential)
  2  dzPhi from cosh(k DZ) instead of sinh(k DZ)
"""
from __future__ import annotations

import numpy as np
from scipy.interpolate import CubicSpline


def _grad(a: np.ndarray, d: float) -> np.ndarray:
    g = np.empty_like(a)
    g[0] = (a[1] - a[0]) / d
    g[-1] = (a[-1] - a[-2]) / d
    g[1:-1] = (a[2:] - a[:-2]) / (2 * d)
    return g


def kinematics(s, eta, u_hat, deta, dphi_hat, t: float, tprev: float, velx=None, velz=None, plant: int = 0) -> dict:
    """kinematic_modul for the setup s (a Kinematics1D); fields [nxout][nz]."""
    k, depth, z = s.k, s.depth, s.z
    eta = np.asarray(eta, float)
    deta = np.asarray(deta, float)
    u_hat = np.asarray(u_hat, complex)
    dphi_hat = np.asarray(dphi_hat, complex)
    sp = lambda x, y: CubicSpline(x, y, bc_type="natural")  # noqa: E731
    s_hm, s_hp = sp(s.htab, s.gam_hmin), sp(s.htab, s.gam_hplus)
    s_dm, s_dp = sp(s.dztab, s.gam_dzmin), sp(s.dztab, s.gam_dzplus)
    H = eta + depth
    ghm, ghp = s_hm(H), s_hp(H)
    cosh_kh = ghm * np.cosh(k * s.hmin) + ghp * np.cosh(k * s.hplus)
    tanh_kh = ghm * np.tanh(k * s.hmin) + ghp * np.tanh(k * s.hplus)
    ks = np.where(k == 0, 1.0, k)
    phihat = np.where(k == 0, 0.0, u_hat.imag / ks) + 1j * np.where(k == 0, 0.0, -u_hat.real / ks)
    if plant == 1:
        phihat = -phihat
    grad_h, grad_d = _grad(H, s.dx), _grad(depth, s.dx)
    nz, i0, i1 = z.size, s.ix0, s.ix0 + s.nxout
    out = {f: np.empty((s.nxout, nz)) for f in ("dtPhi", "dxPhi", "dzPhi")}
    for q, zv in enumerate(z):
        one = bool(s.flat)
        if not one:
            dz = depth + zv
            dz[dz < 0] = 0.0
            dzp, dzm = dz.max(), max(dz.min(), 0.0)
            one = dzm == dzp
        if one:
            ck, sk = np.cosh(k * (zv + depth[0])), np.sinh(k * (zv + depth[0]))
        else:
            gm, gp = s_dm(dz), s_dp(dz)
            gm[dz == dzm], gp[dz == dzm] = 1.0, 0.0
            gm[dz == dzp], gp[dz == dzp] = 0.0, 1.0
            ck = gm * np.cosh(k * dzm) + gp * np.cosh(k * dzp)
            sk = gm * np.sinh(k * dzm) + gp * np.sinh(k * dzp)
        ratio = ck / cosh_kh
        fac = np.fft.ifft(phihat * ratio * (-k * tanh_kh)).real
        dti = np.fft.ifft(dphi_hat * ratio).real
        dzi = np.fft.ifft(phihat * k * (ck if plant == 2 else sk) / cosh_kh).real
        dxi = np.fft.ifft(1j * k * phihat * ratio).real
        mask = (zv > eta) | (zv < -depth)
        dt = np.where(mask, np.nan, dti + deta * fac)
        dxv = np.where(mask, np.nan, dxi + grad_h * fac + grad_d * dzi)
        dzv = np.where(mask, np.nan, dzi)
        out["dtPhi"][:, q] = dt[i0:i1]
        out["dxPhi"][:, q] = dxv[i0:i1]
        out["dzPhi"][:, q] = dzv[i0:i1]
    out["P"] = -out["dtPhi"] - s.grav * z[None, :] - 0.5 * (out["dxPhi"] ** 2 + 2 * out["dzPhi"])
    vx0 = 0.0 if velx is None else np.asarray(velx, float)
    vz0 = 0.0 if velz is None else np.asarray(velz, float)
    out["acc_x"] = (out["dxPhi"] - vx0) / (t - tprev)
    out["acc_z"] = (out["dzPhi"] - vz0) / (t - tprev)
    return out
