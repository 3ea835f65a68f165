"""1-D internal kinematics (kinematic_modul, 1d/kinematic_modul.c:1-247):
host setup and the oracle, on the CPU.

The oracle is pinned by the linear-wave closed forms on a flat bottom at
rest, and each planted misreading of the oracle fails them; the setup's
sinhcosh2IP weights by the identity they are solved from."""
import ctypes as ct
import math

import numpy as np
import pytest

from hawassi_c_amd import _lib
from hawassi_c_amd.kinematics1d import NH, Kinematics1D, kin_levels, sinhcosh2ip
from hawassi_c_amd.setup1d import invers_omega, omega_i
from oracle.kinematics1d_oracle import kinematics


def flat_wave_case(n=256, dx=1.0, h=10.0, m=8, A=0.7, B=0.3):
    """phi = A cos(kx), phi_t = B cos(kx) on a flat bottom at rest, with the
    peak frequency of that k so that the H weights are exact at H = h (a
    node of the H table: min/max_zsurf = -/+0.5)."""
    x = np.arange(n) * dx
    k = 2 * np.pi * np.fft.fftfreq(n, d=dx)
    km = k[m]
    nu = omega_i(km, h)
    depth = np.full(n, h)
    s = Kinematics1D(x, depth, k, nu, -0.5, 0.5, 11, 12, x[40], x[120])
    u = -A * km * np.sin(km * x)
    fields = dict(eta=np.zeros(n), u_hat=np.fft.fft(u), deta=np.zeros(n), dphi_hat=np.fft.fft(B * np.cos(km * x)))
    xo = x[s.ix0:s.ix0 + s.nxout][:, None]
    zz = s.z[None, :]
    ratio = np.cosh(km * (zz + h)) / np.cosh(km * h)
    exact = dict(dxPhi=-A * km * np.sin(km * xo) * ratio,
                 dzPhi=A * km * np.cos(km * xo) * np.sinh(km * (zz + h)) / np.cosh(km * h),
                 dtPhi=B * np.cos(km * xo) * ratio)
    wet = (zz <= 0.0) & (zz >= -h)
    return s, fields, exact, np.broadcast_to(wet, exact["dxPhi"].shape)


def test_sinhcosh2ip_reproduces_cosh_and_sinh_at_the_peak_wavenumber():
    """gamma_- cosh(kap Dmin) + gamma_+ cosh(kap Dplus) = cosh(kap D) and
    likewise for sinh, kap = invers_omega(nu, D) (1d/functions.c:1889-1930);
    the ends give (1, 0) and (0, 1)."""
    D = np.linspace(3.0, 12.0, NH + 1)
    nu = 0.8
    gm, gp = sinhcosh2ip(D, D[0], D[-1], nu, 0.05)
    kap = np.array([invers_omega(nu, d) for d in D])
    for f in (np.cosh, np.sinh):
        np.testing.assert_allclose(gm * f(kap * D[0]) + gp * f(kap * D[-1]), f(kap * D), rtol=1e-12)
    assert abs(gm[0] - 1) < 1e-12 and abs(gp[0]) < 1e-12
    assert abs(gm[-1]) < 1e-12 and abs(gp[-1] - 1) < 1e-12
    # D = 0 takes kap = k[2]
    g0m, g0p = sinhcosh2ip(np.array([0.0]), 0.0, 5.0, nu, 0.05)
    assert abs(g0m[0] - 1) < 1e-12 and abs(g0p[0]) < 1e-12


def test_levels_surface_layer_and_deep_layer():
    """modified_influx's levels (1d/functions.c:3617-3650): nsurf equidistant
    levels over [min_zsurf, max_zsurf], ndeep deeper ones below min_zsurf
    whose spacing grows with depth, sorted ascending."""
    z = kin_levels(-0.05, 0.06, 50, 40, 0.9, 8.0)
    assert z.size == 90 and np.all(np.diff(z) >= 0)
    surf = -0.05 + np.arange(50) * (0.11 / 49)
    assert np.all(np.isin(surf, z))
    deep = np.sort(z[z < -0.05])[::-1]          # from the surface down
    assert deep.size == 40
    steps = -np.diff(deep)
    assert np.all(steps > 0) and np.all(np.diff(steps) > 0)   # sinh decays with depth
    assert deep[0] == pytest.approx(-(0.05 + 0.11 / 49), abs=1e-15)


@pytest.mark.parametrize("plant", [0, 1, 2])
def test_oracle_linear_wave_closed_forms(plant):
    s, f, exact, wet = flat_wave_case()
    out = kinematics(s, f["eta"], f["u_hat"], f["deta"], f["dphi_hat"], 1.0, 0.9, plant=plant)
    for name in ("dtPhi", "dxPhi", "dzPhi"):
        assert np.array_equal(np.isnan(out[name]), ~wet), name
    err = max(np.max(np.abs(out[q][wet] - exact[q][wet])) / np.max(np.abs(exact[q][wet]))
              for q in ("dtPhi", "dxPhi", "dzPhi"))
    if plant == 0:
        assert err < 1e-6, err
        # pressure and accelerations as the reference writes them
        P = -out["dtPhi"] - s.grav * s.z[None, :] - 0.5 * (out["dxPhi"] ** 2 + 2 * out["dzPhi"])
        assert np.array_equal(np.isnan(P), ~wet)
        np.testing.assert_allclose(out["P"][wet], P[wet], rtol=0, atol=0)
        np.testing.assert_allclose(out["acc_x"][wet], out["dxPhi"][wet] / 0.1, rtol=1e-12)
    else:
        assert err > 0.1, err


def test_setup_refuses_dry_nodes_and_the_entry_refuses_eta_outside_the_table():
    n = 256
    x = np.arange(n) * 1.0
    k = 2 * np.pi * np.fft.fftfreq(n)
    d = np.full(n, 10.0)
    d[5] = 0.0
    with pytest.raises(NotImplementedError):
        Kinematics1D(x, d, k, 0.5, -0.5, 0.5, 5, 5, x[10], x[20])
    s, f, _, _ = flat_wave_case()
    lib = _lib.load()
    eta = np.zeros(s.n)
    eta[3] = 0.6                                  # H = 10.6 > Hplus = 10.5
    cfg = s.config()
    dp = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_double))  # noqa: E731
    uh = np.ascontiguousarray(f["u_hat"].view(float))
    out = np.empty(6 * s.nxout * s.z.size)
    rc = lib.hw1d_kinematics(ct.byref(cfg), dp(eta), dp(uh), dp(eta), dp(uh), 1.0, 0.9, None, None, dp(out))
    assert rc == 1                                # HW_ERR_ARG, before any device use
    assert b"outside the table" in lib.hw_last_error()
    assert math.isfinite(s.hmin) and s.flat == 1
