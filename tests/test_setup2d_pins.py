"""Pinning the shared 2-D host setup (no GPU).

The oracle and the device path both read what hawassi_c_amd/setup2d.py
builds (the reference's fun_param_init / fun_operator_setup products), so a
misreading there would be invisible to every parity test.  Each function is
checked here against values fixed by hand from the reference's text or by
the mathematics (closed forms, independent root finding, exact identities),
and each check is shown to fail on a plausible misreading.
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

from hawassi_c_amd import setup2d as S

G = S.GRAV


# --------------------------------------------------------------- freqspace --
def test_freqspace_2d_values():
    """freqspace (2d/functions.c:1810-1832) at L = 1500 m, N = 2048:
    spacing (fmax - fmin)/(N-1) with fmin = -pi (N-1)/L, fmax = -fmin - 2 pi/L,
    i.e. d = 2 pi (N-2) / (L (N-1)), not 2 pi / L; k[1] = 4.185721e-3 by
    hand: (pi/L)(2050 (N-2)/(N-1) - (N-1)) = (pi/1500) 1.9985344; k[0] is
    forced to 1e-10; k[N/2] = fmin."""
    k = S.freqspace(1500.0, 2048)
    assert k[0] == 1e-10
    assert k[1] == pytest.approx(4.185721e-3, rel=2e-7)
    assert k[1024] == pytest.approx(-math.pi * 2047 / 1500, rel=1e-14)
    d = 2 * math.pi * 2046 / (1500 * 2047)
    assert np.allclose(np.diff(k[1:1024]), d, rtol=1e-9)
    assert np.allclose(np.diff(k[1024:]), d, rtol=1e-9)
    # the grid is not antisymmetric: k[n-j] + k[j] = -2 pi / (L (n-1)) for 1 <= j < n/2
    j = np.arange(1, 1024)
    assert np.allclose(k[2048 - j] + k[j], -2 * math.pi / (1500 * 2047), rtol=1e-6, atol=0)
    # a plausible misreading: the textbook FFT grid 2 pi fftfreq(N, dx), dx = L/(N-1)
    k_tb = 2 * math.pi * np.fft.fftfreq(2048, 1500 / 2047)
    assert abs(k_tb[1] - 4.185721e-3) > 1e-6


def test_freqspace_odd_n():
    """Odd N: fmax = -fmin, so d = 2 pi / L; k = [1e-10, d, 2d, ..., then the
    negative half starting at fmin = -pi (N-1)/L]."""
    L, n = 10.0, 5
    k = S.freqspace(L, n)
    d = 2 * math.pi / L
    assert k[0] == 1e-10
    assert k[1] == pytest.approx(d, rel=1e-14) and k[2] == pytest.approx(-2 * d, rel=1e-14)
    assert k[3] == pytest.approx(-d, rel=1e-14) and k[4] == pytest.approx(0.0, abs=1e-15)


# --------------------------------------------------------------- fun_aal2D --
def test_aal2d_radial_cut_point():
    """fun_aal2D (2d/functions.c:1834-1879): idx = floor(Nx/cutfrac) - 1,
    idy = floor(Ny/cutfrac) - 1, keep |k| <= sqrt(kx[idx]^2 + ky[idy]^2).
    With Nx = 32, Ny = 16, cutfrac 4: idx = 7, idy = 3.  On a square-spaced
    grid the point (kx[7], ky[3]) is exactly on the circle and kept; the
    neighbours (kx[8], ky[3]) and (kx[7], ky[4]) are outside."""
    kx = S.freqspace(31.0, 32)      # dx = dy = 1
    ky = S.freqspace(15.0, 16)
    a = S.aal2d(4, kx, ky)
    assert a[3, 7] == 1 and a[3, 8] == 0 and a[4, 7] == 0
    # the negative-k half is cut by |k| too, and since freqspace gives
    # |k[n-j]| = |k[j]| + 2 pi/(L (n-1)) the mirror of the boundary point lies
    # just outside the circle
    assert a[16 - 3, 32 - 7] == 0 and a[16 - 3, 32 - 6] == 1 and a[16 - 2, 32 - 7] == 1
    # the mask is exactly the disk: count of kept points
    R2 = kx[7] ** 2 + ky[3] ** 2
    inside = (kx[None, :] ** 2 + ky[:, None] ** 2) <= R2
    assert np.array_equal(a == 1, inside)
    # misreading "idx = floor(Nx/cutfrac)" (no -1) keeps (kx[8], ky[3])
    R2w = kx[8] ** 2 + ky[4] ** 2
    assert (kx[8] ** 2 + ky[3] ** 2) <= R2w


def test_aal2d_separable_fallback():
    """idx = floor(4/4) - 1 = 0: the separable branch, |kx| <= |kx[0]| = 1e-10
    keeps only column 0, and |ky| <= |ky[idy]| with idy = floor(32/4) - 1 = 7."""
    kx = S.freqspace(3.0, 4)
    ky = S.freqspace(31.0, 32)
    a = S.aal2d(4, kx, ky)
    assert np.all(a[:, 1:] == 0)
    keep_rows = np.abs(ky) <= abs(ky[7])
    assert np.array_equal(a[:, 0] == 1, keep_rows)
    assert keep_rows.sum() == 14     # rows 0..7 and the 6 negative rows with |ky| below ky[7]


# -------------------------------------------------------------- fun_cutoff --
def test_cutoff_keeps_the_reference_precedence():
    """fun_cutoff (2d/functions.c:225-258) computes -b + sqrt(b^2-4a(c-R))/(2a)
    (the division binds to the square root only).  By hand at R = 0.5:
    HS2 (a, b, c) = (0.7289, 0.3229, -0.00058): sqrt(0.10426441 + 1.45949105)
    / 1.4578 - 0.3229 = 0.534901; HS1 (0.8086, 0.2413, 0.0044): 0.555678.
    The textbook root (-b + sqrt(.))/(2a) would give 0.636 / 0.648."""
    assert S.cutoff(0.5, "HS2") == pytest.approx(0.5349009, abs=2e-7)
    assert S.cutoff(0.5, "HS1") == pytest.approx(0.5556783, abs=2e-7)
    assert S.cutoff(1.0) == 1.0 and S.cutoff(0.0) == 0.0
    a, b, c = 0.7289, 0.3229, -0.00058
    textbook = (-b + math.sqrt(b * b - 4 * a * (c - 0.5))) / (2 * a)
    assert abs(textbook - 0.5349009) > 0.05


# -------------------------------------------------- fun_wallchar (user) --
def test_wallchar_user_defined():
    """fun_wallchar user_defined branch (2d/functions.c:264-283): wallchar =
    1 - w cutoff(R), cfSAwall = 1 - w set to 0 wherever it is below 1."""
    w = np.zeros((6, 8))
    w[2:4, 3:6] = 1.0
    w[5, 0] = 0.5                  # a partial value: wallchar scales, cfSAwall still 0
    wc, cfw = S.wallchar_user(w, 0.5)
    co = 0.5349009
    assert np.allclose(wc[2:4, 3:6], 1 - co, atol=1e-6)
    assert wc[5, 0] == pytest.approx(1 - 0.5 * co, abs=1e-6)
    assert wc[0, 0] == 1.0 and cfw[0, 0] == 1.0
    assert np.all(cfw[2:4, 3:6] == 0) and cfw[5, 0] == 0
    assert cfw.sum() == 48 - 7


# ---------------------------------------------- fun_wallchar (rectangles) --
def test_wallchar_rectangles_by_hand():
    """fun_wallchar rectangle branch (2d/functions.c:285-345) on x = 0..15,
    y = 0..9 (unit spacing).  Wall 0 = (4, 2, 8, 5): the fun_Heav2d profiles
    cover 4 <= x < 8 and 2 <= y < 5 (the upper edges are excluded), while
    the damping box cfSAwall = 0 spans the rounded indices 4..8 x 2..5
    inclusive -- one column and one row more than the wall.  Wall 1 =
    (0, 7, 15, 9) starts on the first point and ends on the last one (both
    fun_Heav2d bounds are 0, so both ends are included); it changes wallchar
    but not cfSAwall, which the reference builds from wall 0 alone."""
    x = np.arange(16.0)
    y = np.arange(10.0)
    co = 0.5349009
    wc, cfw = S.wallchar_rect(x, y, [(4.0, 2.0, 8.0, 5.0), (0.0, 7.0, 15.0, 9.0)], 0.5)
    expect = np.ones((10, 16))
    expect[2:5, 4:8] = 1 - co
    expect[7:10, :] = 1 - co
    assert np.allclose(wc, expect, atol=1e-6, rtol=0)
    box = np.ones((10, 16))
    box[2:6, 4:9] = 0.0
    assert np.array_equal(cfw, box)
    # a wall inside [x0, x1] with both edges (a misreading of the Heaviside
    # profiles) would also mark column 8 and row 5
    assert wc[3, 8] == 1.0 and wc[5, 5] == 1.0 and wc[4, 7] < 1.0
    # "influxing" keeps wallchar at 1 (no energy truncation), the damping box stays
    wc2, cfw2 = S.wallchar_rect(x, y, [(4.0, 2.0, 8.0, 5.0)], 0.5, wallmethod="influxing")
    assert np.array_equal(wc2, np.ones((10, 16))) and np.array_equal(cfw2, box)
    # HS1 uses its own cutoff coefficients
    wc3, _ = S.wallchar_rect(x, y, [(4.0, 2.0, 8.0, 5.0)], 0.5, model="HS1")
    assert wc3[3, 5] == pytest.approx(1 - 0.5556783, abs=1e-6)
    # R = 1: fully reflecting, wallchar 0 on the wall
    wc4, _ = S.wallchar_rect(x, y, [(4.0, 2.0, 8.0, 5.0)], 1.0)
    assert wc4[3, 5] == 0.0 and wc4.sum() == 160 - 12


def test_read_wall_dat_stops_as_fscanf_does(tmp_path):
    """wall.dat is read with `if (!fscanf(...)) break` (2d/functions.c:271-
    274): at the end of the file fscanf returns EOF (-1), the loop goes on and
    the remaining points keep wallchar's initial 1, i.e. are wall; a token
    that does not parse returns 0 and ends its row and every later row."""
    f = tmp_path / "wall.dat"
    f.write_text("0 1 0 1\n1 1 0 0\n0 0")
    w = S.read_wall_dat(f, 4, 3)
    assert w.tolist() == [[0, 1, 0, 1], [1, 1, 0, 0], [0, 0, 1, 1]]
    f.write_text("0 1 0 1\n1 x 0 0\n1 1 1 1")
    w = S.read_wall_dat(f, 4, 3)
    assert w.tolist() == [[0, 1, 0, 1], [1, 0, 0, 0], [0, 0, 0, 0]]


# ------------------------------------------------------------- fun_gradient --
def test_gradient_clipped_step():
    """fun_gradient (2d/functions.c:392-461), used on the breaking flux
    (2d/fun_odesolver.c:1001): periodic forward difference clipped to <= 0
    plus periodic backward difference clipped to >= 0.  On a box step of
    height 1 (columns 3..5) with dx = 2: +0.5 at the rising edge (column 3),
    -0.5 at column 5 (the last column of the box), zero elsewhere; the
    centred difference would give +-0.25 on both sides of each edge."""
    f = np.zeros((3, 10))
    f[:, 3:6] = 1.0
    gx, gy = S.gradient_clipped(f, 2.0, 1.0)
    exp = np.zeros(10)
    exp[3], exp[5] = 0.5, -0.5
    assert np.array_equal(gx, np.broadcast_to(exp, (3, 10)))
    assert np.array_equal(gy, np.zeros((3, 10)))
    # periodic wrap on a ramp 0, 1, 2, 3, 4: the rising steps keep their
    # backward difference (+1), column 4 keeps its forward difference to
    # column 0 (-4) plus its backward one (+1), column 0 keeps nothing
    f2 = np.tile(np.arange(5.0), (2, 1))
    gx2, _ = S.gradient_clipped(f2, 1.0, 1.0)
    assert np.array_equal(gx2[0], np.array([0.0, 1, 1, 1, -3]))
    centred = (np.roll(f, -1, axis=1) - np.roll(f, 1, axis=1)) / 4.0
    assert not np.allclose(centred, gx)


# ------------------------------------------------- fun_cfSA2D / fun_fbdy2D --
def _grid():
    x = np.arange(11.0)          # dx = 1, 0..10
    y = np.arange(9.0)           # dy = 1, 0..8
    return x, y


def test_cfsa2d_hand_values():
    """fun_cfSA2D (2d/functions.c:2084-2132) on x = 0..10, y = 0..8 with
    strips Lleft 2, Lright 3, Lbottom 0, Ltop 2: indxl = 2, indxr = 7,
    indyt = 6.  cfX = (1 - cos(pi x / 2))/2 in the left strip (x=1: 0.5),
    1 from x = 2 to 7, 1 - (1 - cos(pi (x - 7)/3))/2 in the right strip
    (x = 8: 0.75, x = 9: 0.25) and 0 at the last point; Lbottom = 0 gives
    cfYB = 1; cfY = 1 - (1 - cos(pi (y - 6)/2))/2 for y > 6 (y = 7: 0.5),
    0 at y = 8."""
    x, y = _grid()
    c = S.cfsa2d(x, y, 2.0, 3.0, 0.0, 2.0)
    cfx = np.array([0, 0.5, 1, 1, 1, 1, 1, 1, 0.75, 0.25, 0.0])
    cfy = np.array([1, 1, 1, 1, 1, 1, 1, 0.5, 0.0])
    assert np.allclose(c, cfy[:, None] * cfx[None, :], atol=1e-15)


def test_fbdy2d_carries_the_strip_length():
    """fun_fbdy2D (2d/functions.c:2026-2082) divides (1 - cfSA) by 0.9 L with
    L a variable that persists across the loops: every point between the
    left and right strips that is not in the bottom/top strip keeps the value
    set at column indxl (Lleft).  That matters inside walls, where cfSA is
    set to 0 (fmin with cfSAwall, :1964-1969): a wall block in the interior
    is damped with 1/(0.9 Lleft) (x10 with walls and model != HS1), not
    1/(0.9 * 7).  Corner points j <= indxl with i > indyt take Lleft too
    (the top-strip rule needs indxl < j < indxr)."""
    x, y = _grid()
    Ll, Lr, Lb, Lt = 2.0, 3.0, 0.0, 1.5
    cf = S.cfsa2d(x, y, Ll, Lr, Lb, Lt)
    wall = np.ones((9, 11))
    wall[3:5, 4:6] = 0.0                      # cfSAwall: a block inside the domain
    cfw = np.minimum(cf, wall)
    fb = S.fbdy2d(x, y, Ll, Lr, Lb, Lt, cfw, wall=True, model="HS2")
    assert np.allclose(fb[3:5, 4:6], 10 / (0.9 * Ll), rtol=1e-14)
    # indyt = closest(y, 8 - 1.5) = 6 (first index of the minimum distance 0.5)
    # top strip, interior columns: L = Ltop
    assert fb[8, 5] == pytest.approx(10 * (1 - cf[8, 5]) / (0.9 * Lt))
    # top strip, left columns: still Lleft
    assert fb[8, 1] == pytest.approx(10 * (1 - cf[8, 1]) / (0.9 * Ll))
    # right strip: Lright
    assert fb[4, 9] == pytest.approx(10 * (1 - cf[4, 9]) / (0.9 * Lr))
    # without walls and for HS1 there is no factor 10
    fb1 = S.fbdy2d(x, y, Ll, Lr, Lb, Lt, cfw, wall=True, model="HS1")
    assert np.allclose(fb1, fb / 10)
    # the misreading "interior L = 7 (the initial value)" changes the wall block
    assert not np.isclose(fb[3, 4], 10 / (0.9 * 7))


def test_fbdy2d_zero_strip_uses_seven():
    """A zero strip length is replaced by 7 when it is the one in force
    (2d/functions.c:2043-2045): with Lleft = 0 the carried value is 7."""
    x, y = _grid()
    cf = S.cfsa2d(x, y, 0.0, 3.0, 0.0, 0.0)
    wall = np.ones((9, 11))
    wall[4, 5] = 0.0
    fb = S.fbdy2d(x, y, 0.0, 3.0, 0.0, 0.0, np.minimum(cf, wall), wall=False)
    assert fb[4, 5] == pytest.approx(1 / (0.9 * 7))


# --------------------------------------------------------------- ChiAdj ----
def test_chi_adjust_line_profile():
    """ChiAdj for a full influx line (2d/functions.c:917-929): 1 - cfv(x; X - dl)
    + cfv(x; X) with cfv = heav(x - Xi) max(sign(x - Xi - dl), (1 - cos((x - Xi)
    pi/dl))/2): 1 at X - dl, 0.5 at X - dl/2, 0 on the line X, 0.5 at X + dl/2,
    1 from X + dl on."""
    x = np.linspace(0.0, 100.0, 101)
    y = np.linspace(0.0, 20.0, 5)
    chi = S.chi_adjust_line(x, y, 50.0, 2.0, 10.0)   # dl = 20
    row = chi[2]
    for xv, cv in ((30, 1.0), (40, 0.5), (50, 0.0), (60, 0.5), (70, 1.0), (0, 1.0), (100, 1.0)):
        assert row[xv] == pytest.approx(cv, abs=1e-14), xv
    assert np.all(chi == chi[0])


# ------------------------------------------- funOprt_DispInterpolation_2p --
def _two_depth_setup(n=32, nk=256):
    """The harbour's depth field (4..20 m) with the wavenumber grid of a
    1.5 km domain at nk points: the reference's 101-point |k| table spans
    [0, max |k|], which must reach the peak wavenumber at the shallowest
    depth (0.125 rad/m at 4 m for Tp = 8 s) or the spline inversion
    extrapolates (the reference's 2p routine has no guard for that, unlike
    its 3p twin, 2d/functions.c:1646-1648)."""
    from hawassi_c_amd import workloads as W
    m = W.harbour(n, t_end=3.0, breaking=False)
    m.kx = S.freqspace(1500.0, nk)[: n] if nk != n else m.kx
    m.ky = S.freqspace(1500.0, nk)[: n] if nk != n else m.ky
    nu = 2 * math.pi / 8.0
    kx = S.freqspace(1500.0, nk)
    o = S.disp_interpolation_2p(m.depth, kx, kx, nu)
    return m, nu, o


def _omega(k, D):
    return math.sqrt(G * k * math.tanh(k * D))


def test_two_depth_weights_fit_phase_speed_and_omega_squared():
    """funOprt_DispInterpolation_2p (2d/functions.c:1534-1625): at every table
    depth D the weights solve g- C(D-) + g+ C(D+) = C(D) and g- w^2(D-) +
    g+ w^2(D+) = w^2(D) at the wavenumber kappa with w(kappa, D) = nu.  kappa
    here comes from an independent root finder (brentq on the exact
    dispersion), the reference's from a natural spline through a 101-point
    table, so they agree to the spline's accuracy."""
    m, nu, o = _two_depth_setup()
    Dm, Dp = o["d_min"], o["d_max"]
    for i in (0, 13, 50, 99, 100):
        D = o["table_depth"][i]
        kap = brentq(lambda k: _omega(k, D) - nu, 1e-9, 10.0, xtol=1e-15)
        C = lambda d: _omega(kap, d) / kap
        M = np.array([[C(Dm), C(Dp)], [_omega(kap, Dm) ** 2, _omega(kap, Dp) ** 2]])
        g = np.linalg.solve(M, [C(D), _omega(kap, D) ** 2])
        assert o["table_gam_min"][i] == pytest.approx(g[0], abs=2e-5), i
        assert o["table_gam_plus"][i] == pytest.approx(g[1], abs=2e-5), i
    # the end points: (1, 0) exactly at D-; near (0, 1) at D+ (the last table
    # depth sits 0.1 dD beyond D+, 2d/functions.c:1553)
    assert o["table_gam_min"][0] == pytest.approx(1.0, abs=1e-13)
    assert o["table_gam_plus"][0] == pytest.approx(0.0, abs=1e-13)
    i_dp = np.argmin(np.abs(o["gam_plus"] - o["gam_plus"].max()))
    assert o["gam_min"].ravel()[i_dp] == pytest.approx(0.0, abs=1e-3)
    assert o["gam_plus"].ravel()[i_dp] == pytest.approx(1.0, abs=1e-3)


def test_two_depth_weights_reproduce_phase_speed_and_omega():
    """The Cramer identity on the grid: at the peak frequency the
    interpolated phase speed and squared frequency equal the exact ones at
    each point's own depth (spline accuracy); a fit of (C^2, w) instead --
    the 1-D Up2IP choice, 1d/functions.c:2210-2218 -- does not reproduce C."""
    m, nu, o = _two_depth_setup()
    Dm, Dp = o["d_min"], o["d_max"]
    worst_c = worst_w = 0.0
    for D, gm, gp in list(zip(m.depth.ravel(), o["gam_min"].ravel(), o["gam_plus"].ravel()))[::37]:
        kap = brentq(lambda k: _omega(k, D) - nu, 1e-9, 10.0, xtol=1e-15)
        C = lambda d: _omega(kap, d) / kap
        worst_c = max(worst_c, abs(gm * C(Dm) + gp * C(Dp) - C(D)) / C(D))
        worst_w = max(worst_w, abs(gm * _omega(kap, Dm) ** 2 + gp * _omega(kap, Dp) ** 2 - nu ** 2) / nu ** 2)
    assert worst_c < 1e-4 and worst_w < 1e-4
    # plausible misreadings fail the identity: the weights swapped, or kappa
    # taken at one reference depth rather than at each point's own depth
    D = 0.5 * (Dm + Dp)
    kap = brentq(lambda k: _omega(k, D) - nu, 1e-9, 10.0, xtol=1e-15)
    C = lambda d, k=kap: _omega(k, d) / k
    i = int(np.argmin(np.abs(o["table_depth"] - D)))
    gm, gp = np.interp(D, o["table_depth"], o["table_gam_min"]), np.interp(D, o["table_depth"], o["table_gam_plus"])
    assert abs(gp * C(Dm) + gm * C(Dp) - C(D)) / C(D) > 1e-2
    kap_p = brentq(lambda k: _omega(k, Dp) - nu, 1e-9, 10.0, xtol=1e-15)
    Mw = np.array([[C(Dm, kap_p), C(Dp, kap_p)], [_omega(kap_p, Dm) ** 2, _omega(kap_p, Dp) ** 2]])
    gw = np.linalg.solve(Mw, [C(D, kap_p), _omega(kap_p, D) ** 2])
    assert abs(gw[0] - o["table_gam_min"][i]) > 1e-2


def test_two_depth_multipliers_closed_form():
    """The spectral multipliers at the two reference depths: Om2dSq = w^2 =
    g |k| tanh(|k| D) with the directional sign squared away, Csqr2d = Cp^2 =
    g tanh(|k| D)/|k| (2d/functions.c:1591-1607, fun_exact_disp_val /
    fun_exact_Cp_val :101-120)."""
    from hawassi_c_amd import workloads as W
    m = W.harbour(16, t_end=3.0, breaking=False)
    o = S.disp_interpolation_2p(m.depth, m.kx, m.ky, 2 * math.pi / 8.0)
    for (i, j) in ((0, 1), (3, 5), (9, 12), (15, 15)):
        kk = math.hypot(m.kx[j], m.ky[i])
        for key, D in (("min", o["d_min"]), ("plus", o["d_max"])):
            assert o["om2_" + key][i, j] == pytest.approx(G * kk * math.tanh(kk * D), rel=1e-13)
            assert o["csq_" + key][i, j] == pytest.approx(G * math.tanh(kk * D) / kk, rel=1e-13)
