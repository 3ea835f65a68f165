"""Pinning the 2-D oracle's AB2 terms, HSS and breaking pieces (no GPU).

The reference ships no 2-D outputs, so each part of the 2-D oracle that the
half-spectrum restatement (tests/halfspec.py) does not pin independently is
checked here against what the mathematics fixes:

* Hamiltonian structure: with every non-conservative term neutral (aal = 1,
  ChiAdj = 1, no damping, source or breaking) the AB2 right-hand side is
  eta_t = dH/dphi, u_t = -grad dH/deta for H = 1/2 g eta^2 + 1/(2g) u.HSS(u)
  + 1/2 eta w (|u|^2 - (L phi)^2), so dH/dt = 0 to rounding -- HS1 flat, two
  and three reference depths; HS2 flat and two-depth with uniform weights.
  (With spatially varying weights the reference's L of the cubic term is
  HSS_{w^2}, not -div HSS_{C^2} grad, and the two differ by O(grad gamma): the
  two-depth HS2 model is Hamiltonian only to that order -- a property of the
  reference, shown below.  States are band-limited to |k| < 0.3 max|k| so no
  product reaches the Nyquist modes, where the spectral L keeps a mode that
  the derivative form drops);
* HSSgen_2depth / _3depth is symmetric: <psi, HSS phi> = <HSS psi, phi>;
* y-uniform reduction: on fields that do not depend on y the 2-D flat AB2
  right-hand side is the 1-D rhs_HsF2 (whose oracle the reference's own gauge
  records pin, tests/test_case1d.py), and the two-depth HS1 one is the 1-D
  two-depth rhs_HsF1;
* breaking: the Hilbert transform of cos is sin, the local wavenumber of a
  plane wave is the centred-difference |k|, the 3x3 peak rule on hand-built
  fields, and the eddy-viscosity ramp B = U/U* - 1 between U* and 2 U*.

Each check is also run with one planted mistake in the oracle
(oracle.set_plant) and must then fail.
"""
import math

import numpy as np
import pytest
from conftest import relerr

from oracle import oracle as O
from hawassi_c_amd import setup2d as S
from hawassi_c_amd import workloads as W

G = S.GRAV


@pytest.fixture(autouse=True)
def _no_plant():
    O.set_plant(0)
    yield
    O.set_plant(0)


def planes(y, n_y, n_x):
    N = n_x * n_y
    return [(y[2 * f * N:(2 * f + 1) * N] + 1j * y[(2 * f + 1) * N:(2 * f + 2) * N]).reshape(n_y, n_x)
            for f in range(3)]


def phys(spec):
    """The reference's inverse: Hermitian projection, backward DFT / N, real part."""
    return np.real(np.fft.ifft2(O.hermitian_project(spec)))


def conservative(m):
    """Make every non-Hamiltonian term neutral."""
    one = np.ones((m.ny, m.nx))
    m.aal, m.chi, m.fbdy = one, one.copy(), 0.0 * one
    m.source_on = False
    m.breaking = None
    return m


def flat_model(n=32):
    m = conservative(W.short_crested(n, t_end=3.0))
    g = W.rng(5)
    jj, ii = np.meshgrid(np.arange(n), np.arange(n))
    # a non-trivial (smooth, band-limited) wall weight enters H2 consistently
    m.wallchar = 0.75 + 0.2 * np.cos(2 * math.pi * jj / n) + 0.05 * np.sin(2 * math.pi * (jj + 2 * ii) / n + g.uniform())
    return m


def two_depth_model(n=32, three=False):
    m = conservative(W.harbour(n, t_end=3.0, breaking=False))
    if three:
        o = S.disp_interpolation_3p(m.depth, m.kx, m.ky, 2 * math.pi / 8.0)
        m.ops = {**m.ops, **{k: o[k] for k in ("gam_min", "gam_mid", "gam_plus", "csq_min", "csq_mid",
                                               "csq_plus", "om2_min", "om2_mid", "om2_plus")}}
    return m


class Hamiltonian:
    """H(eta, u, v) of the AB (HS1 / HS2) model on the reference's grid, with
    the oracle's HSS for a varying bottom and C2d^2 for a flat one."""

    def __init__(self, m, order):
        self.m, self.order = m, order
        self.o = O.Oracle2D(m)

    def csq(self, f):
        m = self.m
        if m.flat:
            return m.c2d ** 2 * np.fft.fft2(f)
        return self.o.hss(f, "csq")

    def __call__(self, y):
        m = self.m
        eh, uh, vh = planes(y, m.ny, m.nx)
        eta, u, v = phys(eh), phys(uh), phys(vh)
        cu, cv = self.csq(u), self.csq(v)
        H = 0.5 * G * np.sum(eta ** 2) + (np.sum(u * phys(cu)) + np.sum(v * phys(cv))) / (2 * G)
        if self.order >= 2:
            KX, KY = m.kx[None, :], m.ky[:, None]
            lphi = phys(-1j * KX * cu / G - 1j * KY * cv / G)
            H += 0.5 * np.sum(eta * m.wallchar * (u * u + v * v - lphi * lphi))
        return H


def dH_dt(m, order, y, t=0.4):
    """dH/dt = DH(y).R(y) by a Richardson combination that is exact for the
    cubic H, and the power of the gravity term alone as scale."""
    m.order = order
    o = O.Oracle2D(m)
    r = o.rhs(t, y)
    H = Hamiltonian(m, order)
    e = 1e-3
    d1 = H(y + e * r) - H(y - e * r)
    d2 = H(y + 2 * e * r) - H(y - 2 * e * r)
    dh = (8 * d1 - d2) / (12 * e)
    eta = phys(planes(y, m.ny, m.nx)[0])
    eta_t = phys(planes(r, m.ny, m.nx)[0])
    return dh, abs(G * np.sum(eta * eta_t))


def uniform_weights(m, a=0.3):
    one = np.ones((m.ny, m.nx))
    m.ops = {**m.ops, "gam_min": a * one, "gam_plus": (1 - a) * one}
    return m


def make(kind):
    if kind == "flat":
        return flat_model()
    if kind == "uniform":
        return uniform_weights(two_depth_model())
    return two_depth_model(three=(kind == "three"))


@pytest.mark.parametrize("kind,order", [("flat", 2), ("flat", 1), ("two", 1), ("three", 1), ("uniform", 2)])
def test_hamiltonian_is_conserved(kind, order):
    m = make(kind)
    y = W.random_state(m, 21, amp=0.4, kfrac=0.3)
    dh, scale = dH_dt(m, order, y)
    assert scale > 1e-3
    assert abs(dh) / scale < 5e-9, (dh, scale)   # kx[0] = 1e-10 leaves a 1e-10-level term


def test_two_depth_hs2_hamiltonian_defect_is_the_weight_gradient():
    """Varying weights leave a residual in dH/dt (the reference's L of the
    cubic term is HSS_{w^2}); it vanishes with uniform weights and shrinks
    with the weights' variation."""
    y = W.random_state(two_depth_model(), 21, amp=0.4, kfrac=0.3)
    res = []
    for a in (1.0, 0.5, 0.0):
        m = two_depth_model()
        gm = m.ops["gam_min"]
        m.ops = {**m.ops, "gam_min": a * gm + (1 - a) * gm.mean(), "gam_plus": a * m.ops["gam_plus"]
                 + (1 - a) * m.ops["gam_plus"].mean()}
        dh, scale = dH_dt(m, 2, y)
        res.append(abs(dh) / scale)
    assert res[0] > 1e-5 and res[1] < 0.7 * res[0] and res[2] < 1e-9


@pytest.mark.parametrize("plant", [1, 2, 6])
@pytest.mark.parametrize("kind", ["flat", "uniform"])
def test_hamiltonian_check_fails_on_planted_mistakes(plant, kind):
    """E without its 1/2, the sign of (L phi)^2 in E, ky in the u equation's
    gravity term: each breaks dH/dt = 0."""
    m = make(kind)
    y = W.random_state(m, 21, amp=0.4, kfrac=0.3)
    O.set_plant(plant)
    dh, scale = dH_dt(m, 2, y)
    assert abs(dh) / scale > 1e-4


@pytest.mark.parametrize("three,which", [(False, "csq"), (False, "om2"), (True, "csq"), (True, "om2")])
def test_hss_is_symmetric(three, which):
    """<psi, HSS phi> = <HSS psi, phi> for HSSgen_2depth / _3depth
    (2d/fun_odesolver.c:500-697): 1/2 sum_j [F(g_j F^-1(L_j phi^)) + L_j F(g_j phi)]
    with real multipliers; the real inverse of the projected spectrum is the
    operator the right-hand side uses."""
    m = two_depth_model(three=three)
    o = O.Oracle2D(m)
    g = W.rng(8)
    psi, phi = g.standard_normal((m.ny, m.nx)), g.standard_normal((m.ny, m.nx))
    a = np.sum(psi * phys(o.hss(phi, which)))
    b = np.sum(phys(o.hss(psi, which)) * phi)
    assert abs(a - b) / abs(a) < 1e-12
    # and it is not trivially symmetric: the plain product gam * L phi is not
    O.set_plant(3)
    a3 = np.sum(psi * phys(O.Oracle2D(m).hss(phi, which)))
    b3 = np.sum(phys(O.Oracle2D(m).hss(psi, which)) * phi)
    assert abs(a3 - b3) / abs(a3) > 1e-3


def test_hss_unit_weight_is_the_multiplier():
    """gam_- = 1, gam_+ = 0: HSS reduces to the single-depth multiplier
    (brute force on a 16 x 16 grid) on the interior columns; on the Nyquist
    column the first half of HSS takes the real part of the inverse, which
    averages the multiplier with its mirror."""
    m = two_depth_model(16)
    one, zero = np.ones((16, 16)), np.zeros((16, 16))
    m.ops = {**m.ops, "gam_min": one, "gam_plus": zero}
    o = O.Oracle2D(m)
    phi = W.rng(9).standard_normal((16, 16))
    got = O.hermitian_project(o.hss(phi, "csq"))
    exp = O.hermitian_project(m.ops["csq_min"] * np.fft.fft2(phi))
    assert relerr(got[:, 1:8], exp[:, 1:8]) < 1e-13
    assert relerr(got[:9, 0], exp[:9, 0]) < 1e-13
    L = m.ops["csq_min"][:, 8]
    Lavg = 0.5 * (L + L[(-np.arange(16)) % 16])
    f = np.fft.fft2(phi)[:, 8]
    assert relerr(got[:, 8], 0.5 * (L + Lavg) * f) < 1e-13


# ----------------------------------------------------- y-uniform reduction --
def _y_uniform_pair(nx=64, ny=4, order=2, two_depth=False, antisym=True):
    """A 2-D model on nx x ny and the 1-D model on nx with the same x grid,
    wavenumbers and operators, every other term neutral.  The 1-D inverse is
    the real part of a full complex backward transform (no projection,
    1d/functions.c:1642-1661) and the 2-D one projects first; the two agree
    for odd multipliers on an antisymmetric wavenumber grid, so the reduction
    uses 2 pi fftfreq (with k[0] = 1e-10 as freqspace sets it); on the
    reference's own grid (k[n-j] + k[j] = -2 pi/(L (n-1))) they differ at
    O(1/n) in the nonlinear terms."""
    from hawassi_c_amd.setup1d import Model1D
    L = 200.0
    x = S.interval_to_vector(0.0, L, nx)
    y = S.interval_to_vector(0.0, 4.0 * (ny - 1), ny)
    if antisym:
        kx = 2 * math.pi * np.fft.fftfreq(nx, x[1] - x[0])
        ky = 2 * math.pi * np.fft.fftfreq(ny, y[1] - y[0])
        kx[0] = ky[0] = 1e-10
    else:
        kx = S.freqspace(L, nx)
        ky = S.freqspace(y[-1] - y[0], ny)
    KX, KY = kx[None, :], ky[:, None]
    D = 2.0
    one = np.ones((ny, nx))
    c2d = S.exact_cp2d(KX, KY, D)
    om = S.exact_disp2d(KX, KY, D)
    m2 = S.Model2D(nx=nx, ny=ny, x=x, y=y, kx=kx, ky=ky, flat=not two_depth, cpeak=1.0, aal=one,
                   fbdy=0 * one, wallchar=one, chi=one, depth=D * one, c2d=c2d, l2d=om * om / G, order=order)
    m1 = Model1D(n=nx, x=x, dx=x[1] - x[0], k=kx.copy(), aal=np.ones(nx), chi=np.ones(nx), damp=np.zeros(nx),
                 depth=np.full(nx, D), flat=not two_depth, csqr=c2d[0] ** 2, lop=(om * om / G)[0],
                 source_on=False, propagation="None", order=order)
    if two_depth:
        gm = 0.6 + 0.3 * np.sin(2 * math.pi * x / L)
        cm, cp = S.exact_cp2d(KX, KY, 1.0) ** 2, S.exact_cp2d(KX, KY, 3.0) ** 2
        om_m, om_p = S.exact_disp2d(KX, KY, 1.0) ** 2, S.exact_disp2d(KX, KY, 3.0) ** 2
        m2.ops = dict(gam_min=gm[None, :] * one, gam_plus=(1 - gm)[None, :] * one, csq_min=cm, csq_plus=cp,
                      om2_min=om_m, om2_plus=om_p)
        m1.ops = dict(gam_min=gm, gam_plus=1 - gm, csq_min=cm[0], csq_plus=cp[0])
    return m2, m1


def _y_uniform_state(nx, ny, seed=4, amp=0.1):
    g = W.rng(seed)
    k = np.fft.fftfreq(nx) * nx
    fields = []
    for f in range(2):
        s = (g.standard_normal(nx) + 1j * g.standard_normal(nx)) * (np.abs(k) < nx // 5)
        p = np.real(np.fft.ifft(s))
        fields.append(p * (amp if f == 0 else 3 * amp) / np.abs(p).max())
    eta, u = fields
    y2 = np.concatenate([np.r_[np.fft.fft2(np.tile(a, (ny, 1))).real.ravel(), np.fft.fft2(np.tile(a, (ny, 1))).imag.ravel()]
                         for a in (eta, u, 0 * u)])
    y1 = np.concatenate([np.r_[np.fft.fft(a).real, np.fft.fft(a).imag] for a in (eta, u)])
    return y2, y1


@pytest.mark.parametrize("order,two_depth", [(2, False), (1, False), (1, True)])
def test_y_uniform_reduces_to_1d(order, two_depth):
    nx, ny = 64, 4
    m2, m1 = _y_uniform_pair(nx, ny, order, two_depth)
    y2, y1 = _y_uniform_state(nx, ny)
    r2 = O.Oracle2D(m2).rhs(0.3, y2)
    r1 = O.Oracle1D(m1).rhs(0.3, y1)
    N = nx * ny
    h = nx // 2 + 1
    for f2, f1 in ((0, 0), (1, 1), (2, 2), (3, 3)):      # Re/Im of eta^_t, u^_t
        a = r2[f2 * N: f2 * N + nx][:h] / ny             # row ky = 0, stored columns
        b = r1[f1 * nx: (f1 + 1) * nx][:h]
        assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max(), (f2, np.abs(a - b).max(), np.abs(b).max())
    # every other row of the 2-D derivative vanishes for y-uniform data
    for f2 in range(4):
        assert np.abs(r2[f2 * N + nx: (f2 + 1) * N]).max() <= 1e-9 * np.abs(r2[f2 * N: f2 * N + nx]).max()


def test_y_uniform_on_the_reference_grid_differs_at_order_one_over_n():
    """With freqspace's grid the 1-D real-part inverse and the 2-D projected
    one differ in the nonlinear terms only, by O(1/n) of those terms."""
    nx, ny = 64, 4
    m2, m1 = _y_uniform_pair(nx, ny, 2, False, antisym=False)
    y2, y1 = _y_uniform_state(nx, ny)
    d2 = O.Oracle2D(m2).rhs(0.3, y2)[: nx] / ny
    d1 = O.Oracle1D(m1).rhs(0.3, y1)[: nx]
    m2.order = m1.order = 1
    l2 = O.Oracle2D(m2).rhs(0.3, y2)[: nx] / ny
    l1 = O.Oracle1D(m1).rhs(0.3, y1)[: nx]
    h = nx // 2 + 1
    assert np.abs(l2 - l1)[:h].max() < 1e-12 * np.abs(l1).max()     # linear parts identical
    dn = np.abs((d2 - l2) - (d1 - l1))[:h].max() / np.abs(d1 - l1)[:h].max()
    assert 1e-4 < dn < 4.0 / nx


def test_y_uniform_reduction_fails_on_planted_mistake():
    nx, ny = 64, 4
    m2, m1 = _y_uniform_pair(nx, ny, 2, False)
    y2, y1 = _y_uniform_state(nx, ny)
    O.set_plant(6)
    r2 = O.Oracle2D(m2).rhs(0.3, y2)
    r1 = O.Oracle1D(m1).rhs(0.3, y1)
    a = r2[2 * nx * ny: 2 * nx * ny + nx] / ny
    b = r1[2 * nx: 3 * nx]
    assert np.abs(a - b).max() > 1e-3 * np.abs(b).max()


# ----------------------------------------------------------------- breaking --
def breaking_model(n=64):
    m = W.harbour(n, t_end=3.0)
    m.source_on = False
    return m


def mode_state(m, p, q, amp):
    """eta = amp cos(2 pi (q x_j/nx + p y_i/ny)) as a 6-plane state."""
    jj, ii = np.meshgrid(np.arange(m.nx), np.arange(m.ny))
    eta = amp * np.cos(2 * math.pi * (q * jj / m.nx + p * ii / m.ny))
    z = np.zeros((m.ny, m.nx))
    return np.concatenate([np.r_[np.fft.fft2(a).real.ravel(), np.fft.fft2(a).imag.ravel()] for a in (eta, z, z)]), eta


@pytest.mark.parametrize("p,q", [(0, 3), (2, 5), (5, 2), (-3, 7)])
def test_hilbert_of_cos_is_sin_and_local_wavenumber(p, q):
    """breaking_criterion (2d/breaking.c:248-281): H eta = F^-1(P(-i
    sign(kx + ky) eta^)) maps cos(k.x) to sign(kx + ky) sin(k.x) for the
    stored (kx >= 0) mode, and the local wavenumber
    |(eta grad H - H grad eta)/(eta^2 + H^2)| of a plane wave with centred
    differences is sqrt((sin(kx dx)/dx)^2 + (sin(ky dy)/dy)^2) at interior points."""
    m = breaking_model(32)
    y, eta = mode_state(m, p, q, 0.7)
    o = O.Oracle2D(m)
    e, Hr, K = o.break_diag(y)
    assert np.abs(e - eta).max() < 1e-13
    jj, ii = np.meshgrid(np.arange(m.nx), np.arange(m.ny))
    th = 2 * math.pi * (q * jj / m.nx + p * ii / m.ny)
    ky_i = p % m.ny
    sgn = 1.0 if (m.kx[q] + m.ky[ky_i]) >= 0 else -1.0
    assert np.abs(Hr - sgn * 0.7 * np.sin(th)).max() < 1e-13
    dx, dy = m.x[1] - m.x[0], m.y[1] - m.y[0]
    kexp = math.hypot(math.sin(2 * math.pi * q / m.nx) / dx, math.sin(2 * math.pi * p / m.ny) / dy)
    assert np.abs(K[1:-1, 1:-1] - kexp).max() < 1e-10 * kexp
    O.set_plant(4)
    _, Hr4, _ = O.Oracle2D(m).break_diag(y)
    assert np.abs(Hr4 - sgn * 0.7 * np.sin(th)).max() > 0.5


def test_peak_rule_hand_built():
    """fun_peaks (2d/breaking.c:55-133): a point is a peak when no neighbour
    exceeds it by more than 0.001 (3x3 inside, 2x3 on the sides, corners never)
    and it is above the threshold; plateaus give one peak per point; peaks are
    listed interior first (x-major), then the sides x = 0, x = Nx-1, y = 0,
    y = Ny-1."""
    nx, ny = 12, 9
    ec = np.zeros((ny, nx))
    ec[4, 5] = 1.0                       # a clean interior peak
    ec[2, 9] = 0.8
    ec[2, 10] = 0.8005                   # a neighbour higher by 5e-4: both are peaks
    ec[6, 0] = 0.9                       # a side peak on x = 0
    ec[0, 0] = 5.0                       # a corner is never a peak ...
    ec[8, 3] = 0.7                       # ... a side peak on y = Ny-1
    ec[5, 5] = 0.995                     # 5e-3 below the interior peak: not a peak
    got = O.find_peaks(ec, 0.5)
    assert got == [(5, 4), (9, 2), (10, 2), (0, 6), (3, 8)]
    assert (0, 0) not in got
    # the threshold is strict
    assert O.find_peaks(ec, 0.9) == [(5, 4)]
    # without the 0.001 tolerance (9, 2) is no peak
    O.set_plant(7)
    assert (9, 2) not in O.find_peaks(ec, 0.5)


def _crest_state(m):
    """A y-uniform ridge of elevation (a Gaussian of 20 m width centred on a
    grid column near x = 750 m) with a fluid-velocity ridge three cells
    ahead of it, so that the speed grows across the crest's search sector:
    U = U0 exp(-((x - xc - 3 dx)/w)^2), U_I = U(xc) = U0 e^{-9 dx^2/w^2}."""
    jc = m.nx // 2
    xc, dx = m.x[jc], m.x[1] - m.x[0]
    x = m.x[None, :] * np.ones((m.ny, 1))
    w = 20.0
    eta = 1.5 * np.exp(-((x - xc) / w) ** 2)
    u = 30.0 * np.exp(-((x - xc - 3 * dx) / w) ** 2)
    v = 0.01 * np.ones_like(u)
    return np.concatenate([np.r_[np.fft.fft2(a).real.ravel(), np.fft.fft2(a).imag.ravel()] for a in (eta, u, v)])


def test_breaking_ramp_between_ustar_and_twice_ustar():
    """At the first detection (flagbr 1, t = tbreak) U* = U_I, the crest's
    speed, and funBr_BreakNodes_char (2d/breaking.c:456-503) sets B = 1 where
    U >= 2 U*, 0 where U <= U*, U/U* - 1 between, on the crest's sector
    where U >= U_F = TC U_I and eta >= 0."""
    m = breaking_model(256)
    o = O.Oracle2D(m)
    y = _crest_state(m)
    o.set_tprev(0.0)
    o.rhs(0.5, y)
    assert o.flagbr() == 2
    B = o.B()
    eta, u, v = (phys(s) for s in planes(y, m.ny, m.nx))
    U = np.hypot(u, v)
    on = B > 0
    assert on.sum() > 10
    # the crest speed U_I: B = U/U_I - 1 on the ramp nodes fixes one U_I for all
    ramp = on & (B < 1)
    ui = U[ramp] / (1 + B[ramp])
    assert np.ptp(ui) < 1e-12 * ui.mean()
    ui = ui.mean()
    assert (B == 1).any() and np.all(U[B == 1] >= 2 * ui * (1 - 1e-14))
    assert np.all(U[on] > ui) and np.all(eta[on] >= 0)
    # the nodes lie within the search radius ahead of the crest column
    jj = np.nonzero(on)[1]
    dxg = m.x[1] - m.x[0]
    assert jj.min() > m.nx // 2 and (jj.max() - m.nx // 2) * dxg <= m.breaking["Rsearch"]
    # U_I is the speed at one of the peaks of eta * Char
    peaks = O.find_peaks(eta * m.breaking["char"], m.breaking["MaxEtaInit"])
    assert any(abs(U[py, px] - ui) < 1e-12 * ui for px, py in peaks)
    O.set_plant(5)
    o5 = O.Oracle2D(m)
    o5.set_tprev(0.0)
    o5.rhs(0.5, y)
    B5 = o5.B()
    r5 = (B5 > 0) & (B5 < 1)
    assert not r5.any() or np.ptp(U[r5] / (1 + B5[r5])) > 1e-3
