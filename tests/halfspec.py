"""Numpy model of the device pipeline's half-spectrum formulation (test aid).

The HIP kernels evaluate the reference's full-grid operator on the half
spectrum (kx = 0..nx/2).  This module restates that formulation with numpy
FFTs so tests can check, on the CPU, that it equals the oracle's full-grid
result on the stored columns:

* inverse transforms use the reference's projection (column kx=0 rebuilt from
  its rows ky < ny/2; imaginary parts at the four self-conjugate points
  zeroed; the Nyquist column's Hermitian part kept by taking real parts);
* fun_adjust keeps the imaginary part of its inverse, which only the Nyquist
  column's anti-Hermitian part produces: (-1)^x s(y) with s the imaginary
  part of that column's y-inverse;
* linear terms are summed before transforming (same operator, other rounding).
"""
from __future__ import annotations

import numpy as np

G = 9.80665


def half(a_full: np.ndarray) -> np.ndarray:
    """[ny, nx] full grid -> [ny, nx/2+1] stored half."""
    return a_full[:, : a_full.shape[1] // 2 + 1]


def c2r(S: np.ndarray, nx: int):
    """Inverse with projection semantics.  S: [ny, nx/2+1].  Returns the
    physical field (normalised by nx*ny) and s(y) (Nyquist imag / (nx*ny))."""
    ny = S.shape[0]
    X = S.astype(complex).copy()
    X[ny // 2 + 1:, 0] = np.conj(X[1: ny // 2, 0][::-1])
    for r in (0, ny // 2):
        X[r, 0] = X[r, 0].real
        X[r, nx // 2] = X[r, nx // 2].real
    W = np.fft.ifft(X, axis=0) * ny
    s = W[:, nx // 2].imag / (nx * ny)
    phys = np.fft.irfft(W, n=nx, axis=1) * nx / (nx * ny)
    return phys, s


def r2c(f: np.ndarray) -> np.ndarray:
    return np.fft.rfft2(f)


def project_eta_half(pre: np.ndarray) -> np.ndarray:
    """The stored-half effect of fun_sbreaking's in-place projection of
    delphiH (2d/fun_odesolver.c:985-1019 via fftwcomplex_2d_sym), applied to
    the pre-mask eta derivative.  The source and damping terms are transforms
    of real fields, so they are already Hermitian along column 0 and real at
    the four self-conjugate bins: projecting the sum equals projecting delphiH."""
    ny, nh = pre.shape
    out = pre.copy()
    out[ny // 2 + 1:, 0] = np.conj(pre[1: ny // 2, 0][::-1])
    for r in (0, ny // 2):
        for c in (0, nh - 1):
            out[r, c] = out[r, c].real
    return out


def rhs_2db(m, t: float, Z: np.ndarray, src_vals=None, project_eta: bool = False) -> np.ndarray:
    """Half-spectrum 2DB RHS; Z is the full 6-plane state; returns the three
    derivative half spectra stacked [3, ny, nx/2+1].  project_eta: the call
    ran breaking (tprev < t), which projects the eta derivative."""
    nx, ny = m.nx, m.ny
    N = nx * ny
    pl = Z.reshape(6, ny, nx)
    eh, uh, vh = [half(pl[2 * f] + 1j * pl[2 * f + 1]) for f in range(3)]
    kx = m.kx[: nx // 2 + 1][None, :]
    ky = m.ky[:, None]
    o = m.ops
    cm, cp = half(o["csq_min"]), half(o["csq_plus"])
    om, op = half(o["om2_min"]), half(o["om2_plus"])
    gm, gp = o["gam_min"], o["gam_plus"]
    w, chi, fb7 = m.wallchar, m.chi, 7 * m.cpeak * m.fbdy
    aal = half(m.aal)

    def hss(phi, phihat, Lm, Lp):
        am, _ = c2r(Lm * phihat, nx)
        ap, _ = c2r(Lp * phihat, nx)
        return 0.5 * (r2c(gm * am + gp * ap) + Lm * r2c(gm * phi) + Lp * r2c(gp * phi))

    eta, _ = c2r(eh, nx)
    u, _ = c2r(uh, nx)
    v, _ = c2r(vh, nx)
    duH1 = -1j * kx * hss(u, uh, cm, cp) / G
    dvH1 = -1j * ky * hss(v, vh, cm, cp) / G
    M0u, _ = c2r(duH1, nx)
    M0v, _ = c2r(dvH1, nx)
    pu, pv = eta * M0u * w, eta * M0v * w
    E = 0.5 * (u * u + v * v - (M0u + M0v) ** 2) * w
    dH2 = (-1j * kx * r2c(eta * u * w) - hss(pu, r2c(pu), om, op) / G
           - 1j * ky * r2c(eta * v * w) - hss(pv, r2c(pv), om, op) / G)
    Eh = r2c(E)
    sign = (-1.0) ** np.arange(nx)[None, :]

    def adj_phys(A):
        a, s = c2r(A, nx)
        return chi * a, chi * sign * s[:, None]

    aeta, seta = adj_phys(dH2)
    ax, sx = adj_phys(1j * kx * Eh)
    ay, sy = adj_phys(1j * ky * Eh)
    src = np.zeros((ny, nx))
    if m.source_on:
        from hawassi_c_amd.setup2d import spline_coeffs, spline_eval
        sv = spline_eval(m.t_sig, m.sig, spline_coeffs(m.t_sig, m.sig), t)
        sk = spline_eval(m.t_sig, m.skew_sig, spline_coeffs(m.t_sig, m.skew_sig), t)
        sv = sv[:, None] if m.vertical else sv[None, :]
        sk = sk[:, None] if m.vertical else sk[None, :]
        if m.propagation == 2:
            src = 2 * sv * m.new_gam
        else:
            sgn = -1.0 if m.propagation == 0 else 1.0
            src = sv * m.new_gam + sgn * sk * m.skew_gam
    Zeta = r2c(src - fb7 * eta + aeta) + 1j * r2c(seta)
    Zu = r2c(-ax - fb7 * u) - 1j * r2c(sx)
    Zv = r2c(-ay - fb7 * v) - 1j * r2c(sy)
    pre = duH1 + dvH1 + Zeta
    de = aal * (project_eta_half(pre) if project_eta else pre)
    du = aal * (-1j * G * kx * eh + Zu)
    dv = aal * (-1j * G * ky * eh + Zv)
    return np.stack([de, du, dv])


def oracle_half(dZ: np.ndarray, nx: int, ny: int) -> np.ndarray:
    pl = dZ.reshape(6, ny, nx)
    return np.stack([half(pl[2 * f] + 1j * pl[2 * f + 1]) for f in range(3)])
