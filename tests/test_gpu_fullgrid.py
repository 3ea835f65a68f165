"""The full-grid 2-D path (solver2dg.h) on the device against the oracle.

This path keeps the reference's full complex state and transforms, so the
comparison runs over every one of the 6 nx ny values, the columns kx > nx/2
included (the fused path returns their Hermitian mirror there instead).
Grids are sides of 2, 3 and 5 factors (the reference's 300x250, 600x300 and
2000x2 cases are such grids; FFTW plans any length).  Bars: FP64 RHS relL2 <=
1e-11 over all values, FP32 <= 3e-5; trajectories FP64 <= 1e-9 with equal step
counts.
"""
import sys

import numpy as np
import pytest
from conftest import relerr

from hawassi_c_amd import HW_PREC_FP32, HW_PREC_FP64
from hawassi_c_amd import _lib
from hawassi_c_amd import workloads as W
from hawassi_c_amd.model2d import Hawassi2D
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def case(name):
    if name == "harbour90x60":          # two depths, source, breaking model, walls (wallchar)
        return W.harbour(90, ny=60, t_end=8.0)
    if name == "flat60x50":             # flat bottom, no walls
        return W.short_crested(60, ny=50, t_end=8.0)
    if name == "flatwall60x50":         # flat bottom, walls of reflection 0.5
        return W.short_crested(60, ny=50, t_end=8.0, wall_refl=0.5)
    if name == "flatwall1_60x50":       # flat bottom, walls of reflection 1
        return W.short_crested(60, ny=50, t_end=8.0, wall_refl=1.0)
    if name == "harbour64":             # a power of two, on the full grid
        return W.harbour(64, t_end=8.0)
    raise KeyError(name)


@pytest.mark.parametrize("nx,ny", [(60, 50), (300, 250), (2000, 2), (64, 32), (90, 36)])
@pytest.mark.parametrize("prec,tol", [(HW_PREC_FP64, 1e-13), (HW_PREC_FP32, 2e-6)])
def test_full_transforms_match_numpy(nx, ny, prec, tol):
    say("full transforms", nx, ny, prec)
    lib = _lib.load()
    assert lib.hw_supported_full(nx, ny, prec) == 1
    g = np.random.default_rng(nx * 7 + ny)
    f = g.standard_normal((ny, nx))
    re, im = np.empty((ny, nx)), np.empty((ny, nx))
    _lib.check(lib.hw_fft2_forward(nx, ny, prec, _lib.ptr(f), _lib.ptr(re), _lib.ptr(im)), "fwd")
    ref = np.fft.fft2(f)
    assert relerr(re + 1j * im, ref) < tol
    back = np.empty((ny, nx))
    _lib.check(lib.hw_fft2_inverse(nx, ny, prec, _lib.ptr(np.ascontiguousarray(ref.real)),
                                   _lib.ptr(np.ascontiguousarray(ref.imag)), _lib.ptr(back), None), "inv")
    assert relerr(back, f) < tol


def test_full_sizes_refused_beyond_2_3_5():
    lib = _lib.load()
    assert lib.hw_supported_full(70, 64, HW_PREC_FP64) == 0      # 7 is a factor
    assert lib.hw_supported_full(8192, 2, HW_PREC_FP64) == 0     # beyond the FP64 LDS transform
    assert lib.hw_supported_full(8192, 2, HW_PREC_FP32) == 1


@pytest.mark.parametrize("name", ["harbour90x60", "flat60x50", "flatwall60x50", "flatwall1_60x50", "harbour64"])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("prec,tol", [(HW_PREC_FP64, 1e-11), (HW_PREC_FP32, 3e-5)])
def test_full_rhs_parity_every_value(name, order, prec, tol):
    m = case(name)
    m.order = order
    say("full rhs", name, "order", order, "prec", prec)
    y = W.random_state(m, 3, kfrac=0.6)
    o = O.Oracle2D(m)
    d = Hawassi2D(m, prec, full_grid=True)
    assert d.full_grid
    # tprev >= t: breaking gated (the sequence test below runs it)
    o.set_tprev(5.0)
    d.set_tprev(5.0)
    ref = o.rhs(1.5, y)
    got = d.rhs(1.5, y)
    e = relerr(got, ref)
    say("   relL2 over all 6 nx ny values %.2e" % e)
    assert e < tol
    # the redundant half on its own: the reference's values, not a mirror
    nx = m.nx
    up = np.concatenate([r.reshape(m.ny, nx)[:, nx // 2 + 1:].ravel() for r in ref.reshape(6, -1)])
    gp = np.concatenate([r.reshape(m.ny, nx)[:, nx // 2 + 1:].ravel() for r in got.reshape(6, -1)])
    assert relerr(gp, up) < tol


def test_full_grid_power_of_two_equals_fused_on_the_stored_half():
    """On a power-of-two grid both paths evaluate the same operator: they
    agree on the columns kx <= nx/2; in the columns above, the full-grid path
    gives the reference's values and the fused path their mirror."""
    m = case("harbour64")
    y = W.random_state(m, 5, kfrac=0.6)
    o = O.Oracle2D(m)
    full = Hawassi2D(m, HW_PREC_FP64, full_grid=True)
    fused = Hawassi2D(m, HW_PREC_FP64)
    assert full.full_grid and not fused.full_grid
    for dd in (o, full, fused):
        dd.set_tprev(5.0)
    ref = o.rhs(1.5, y).reshape(6, 64, 64)
    a = full.rhs(1.5, y).reshape(6, 64, 64)
    b = fused.rhs(1.5, y).reshape(6, 64, 64)
    lo = np.s_[:, :, :33]
    hi = np.s_[:, :, 33:]
    say("stored half: full", relerr(a[lo], ref[lo]), "fused", relerr(b[lo], ref[lo]))
    say("redundant half: full", relerr(a[hi], ref[hi]), "fused (mirror)", relerr(b[hi], ref[hi]))
    assert relerr(a[lo], ref[lo]) < 1e-11 and relerr(b[lo], ref[lo]) < 1e-11
    assert relerr(a[hi], ref[hi]) < 1e-11
    assert relerr(b[hi], ref[hi]) > 1e-9     # the mirror differs from the reference's own values


@pytest.mark.parametrize("name", ["harbour90x60", "flatwall60x50"])
def test_full_trajectory_fp64(name):
    m = case(name)
    m.eps_abs = m.eps_rel = 1e-6
    m.breaking = None
    say("full trajectory", name)
    y0 = W.random_state(m, 7, kfrac=0.4)
    o = O.Oracle2D(m)
    d = Hawassi2D(m, HW_PREC_FP64, full_grid=True)
    y = y0.copy()
    d.set_state(y0)
    rc, t_o, h_o, so = o.integrate(0.0, 0.5, 0.05, y, m.eps_abs, m.eps_rel)
    t_d, h_d = d.integrate(0.0, 0.5, 0.05)
    st = d.stats()
    say("  oracle steps", so, "device", st["steps_accepted"], st["steps_rejected"], "h", h_o, h_d)
    assert rc == 0 and t_o == t_d == 0.5
    assert (st["steps_accepted"], st["steps_rejected"]) == tuple(so)
    assert abs(h_d - h_o) <= 1e-9 * abs(h_o)
    e = relerr(d.get_state(), y)
    say("  state relL2", e)
    assert e < 1e-9


@pytest.mark.parametrize("nx,ny,kbc", [(90, 60, 0.2), (150, 120, 0.2)])
def test_full_breaking_sequence_fp64(nx, ny, kbc):
    say("full breaking sequence", nx, ny, kbc)
    m = W.harbour(nx, ny=ny, t_end=8.0, kbc=kbc)
    y = W.random_state(m, 7, amp=2.0, kfrac=0.15)
    o = O.Oracle2D(m)
    d = Hawassi2D(m, HW_PREC_FP64, full_grid=True)
    o.set_tprev(0.0)
    d.set_tprev(0.0)
    active = 0
    for t in (1.0, 1.05, 1.1, 1.3, 2.0, 3.5):
        ref = o.rhs(t, y)
        got = d.rhs(t, y)
        Bo = o.B()
        Bd, fd = d.breaking_state()
        say(f"  t={t} flagbr oracle={o.flagbr()} device={fd} B>0 {int((Bo > 0).sum())}/{int((Bd > 0).sum())}")
        assert fd == o.flagbr()
        assert np.abs(Bd - Bo).max() < 1e-12
        active += int((Bo > 0).sum() > 0)
        e = relerr(got, ref)
        say("   relL2", e)
        assert e < 1e-10
    assert active > 0, "the state did not trigger breaking; the test would be vacuous"
    o.set_tprev(10.0)
    d.set_tprev(10.0)
    assert relerr(d.rhs(4.0, y), o.rhs(4.0, y)) < 1e-10


def test_full_fields_match_oracle_inverse():
    m = case("harbour90x60")
    y = W.random_state(m, 11, kfrac=0.5)
    d = Hawassi2D(m, HW_PREC_FP64, full_grid=True)
    d.set_state(y)
    eta, u, v = d.fields()
    pl = y.reshape(6, m.ny, m.nx)
    for f, got in enumerate((eta, u, v)):
        ref = O.ifft_c2r(pl[2 * f] + 1j * pl[2 * f + 1])
        assert relerr(got, ref) < 1e-12
    # the full state round-trips exactly (no mirror is imposed on it)
    assert np.array_equal(d.get_state(), y)
