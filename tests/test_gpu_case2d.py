"""Reference cases set up from their input.txt (hawassi_c_amd.case2d) and run
on the device: parity with the oracle on each case's own setup, and the wave
the area source generates (the influx record reproduced downstream)."""
import math
import shutil
import sys
from pathlib import Path

import numpy as np
import pytest
from conftest import relerr

import halfspec as HS
from hawassi_c_amd import HW_PREC_FP32, HW_PREC_FP64
from hawassi_c_amd import case2d as C
from hawassi_c_amd import workloads as W
from hawassi_c_amd.model2d import Hawassi2D
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).parent / "golden" / "input2d"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def case_flat_harmonic(tmp_path):
    d = tmp_path / "2dF001harm"
    d.mkdir()
    shutil.copy(GOLD / "2dF001harm.txt", d / "input.txt")
    return C.model_from_input(str(d))


def case_two_depth(tmp_path):
    """2dB001HarmRefr's input.txt with a synthetic sloping bath.dat."""
    d = tmp_path / "2dB001HarmRefr"
    d.mkdir()
    shutil.copy(GOLD / "2dB001HarmRefr.txt", d / "input.txt")
    nx, ny = 512, 256
    xx = np.linspace(0, 3000, nx)
    yy = np.linspace(-250, 1500, ny)
    depth = 30.0 - 20.0 * xx[None, :] / 3000.0 + 2.0 * np.sin(yy[:, None] / 200.0)
    (d / "bath.dat").write_text("\n".join(" ".join("%.7e" % v for v in row) for row in depth))
    return C.model_from_input(str(d))


def case_user_influx(tmp_path):
    """2dF001Irreg's input.txt (HS1, user_defined influx) with a synthetic
    influx.dat: two harmonics along the 64-point line at x = 0."""
    d = tmp_path / "2dF001Irreg"
    d.mkdir()
    shutil.copy(GOLD / "2dF001Irreg.txt", d / "input.txt")
    nt, ny = 800, 64
    t = np.arange(nt) * 0.5
    y = np.linspace(0, 400, ny)
    rows = ["0 " + " ".join("%f" % 0.0 for _ in range(ny)), "0 " + " ".join("%f" % v for v in y)]
    for j in range(nt):
        e = 0.5 * np.cos(2 * math.pi * t[j] / 8.0) + 0.2 * np.sin(2 * math.pi * t[j] / 5.0 + y / 100.0)
        rows.append("%f " % t[j] + " ".join("%f" % v for v in e))
    (d / "influx.dat").write_text("\n".join(rows))
    return C.model_from_input(str(d))


def case_jonswap(tmp_path):
    """2dF001Irreg's input.txt with `Influx_signal = jonswap` (gamma 1, s 20):
    the line record generated by jonswap_influx, seeded."""
    d = tmp_path / "2dF001Irreg_jonswap"
    d.mkdir()
    txt = (GOLD / "2dF001Irreg.txt").read_text().replace("Influx_signal   = user_defined", "Influx_signal   = jonswap")
    (d / "input.txt").write_text(txt)
    return C.model_from_input(str(d), jonswap_seed=11)


@pytest.mark.parametrize("which", ["flat_harmonic", "two_depth", "user_influx", "jonswap"])
def test_case_rhs_parity(tmp_path, which):
    m = {"flat_harmonic": case_flat_harmonic, "two_depth": case_two_depth, "user_influx": case_user_influx,
         "jonswap": case_jonswap}[which](tmp_path)
    say("case rhs", which, m.nx, m.ny, "order", m.order, "flat", m.flat)
    y = W.random_state(m, 5, amp=0.5, kfrac=0.3)
    ref = HS.oracle_half(O.Oracle2D(m).rhs(37.3, y), m.nx, m.ny)
    for prec, tol in ((HW_PREC_FP64, 1e-11), (HW_PREC_FP32, 3e-5)):
        dev = Hawassi2D(m, prec)
        got = HS.oracle_half(dev.rhs(37.3, y), m.nx, m.ny)
        dev.close()
        e = [relerr(got[f], ref[f]) for f in range(3)]
        say("  prec", prec, "relL2", e)
        assert max(e) < tol


def test_case_trajectory_flat_harmonic(tmp_path):
    """From rest, influx on, to t = 20 s at eps 1e-10 (FP64) against the oracle."""
    m = case_flat_harmonic(tmp_path)
    m.eps_abs = m.eps_rel = 1e-10
    n = m.nx * m.ny
    y = np.zeros(6 * n)
    o = O.Oracle2D(m)
    dev = Hawassi2D(m, HW_PREC_FP64)
    dev.set_state(y)
    t_o, h_o, t_d, h_d = 0.0, 0.05, 0.0, 0.05
    for t1 in (10.0, 20.0):
        rc, t_o, h_o, _ = o.integrate(t_o, t1, h_o, y, m.eps_abs, m.eps_rel)
        assert rc == 0
        t_d, h_d = dev.integrate(t_d, t1, h_d)
    got = HS.oracle_half(dev.get_state(), m.nx, m.ny)
    ref = HS.oracle_half(y, m.nx, m.ny)
    dev.close()
    e = [relerr(got[f], ref[f]) for f in range(3)]
    say("  trajectory relL2", e)
    assert max(e) < 1e-8


def test_case_generated_wave_height(tmp_path):
    """2dF001harm (harmonic influx, Hs = 2 m, Tp = 12 s, 20 m depth, HS3):
    after the ramp the wave height downstream of the influx line is the
    record's, 2 m, to 10 %."""
    m = case_flat_harmonic(tmp_path)
    dev = Hawassi2D(m, HW_PREC_FP32)
    dev.set_state(np.zeros(6 * m.nx * m.ny))
    ix = int(np.argmin(np.abs(m.x - 500.0)))
    iy = m.ny // 2
    t, h = 0.0, 0.05
    t, h = dev.integrate(t, 80.0, h)
    series = []
    for t1 in np.arange(80.5, 104.01, 0.5):
        t, h = dev.integrate(t, float(t1), h)
        eta, _, _ = dev.fields()
        series.append(float(eta[iy, ix]))
    dev.close()
    H = max(series) - min(series)
    say("  wave height at x = %.1f m: %.4f m (record Hs %.2f)" % (m.x[ix], H, m.meta["hs"]))
    assert abs(H - 2.0) < 0.2


# ---- the reference cases on grids that are not powers of two: the full-grid path
CASE2D = Path(__file__).parent / "golden" / "case2d"
REF_CASES = ["2dBBr001_HarmBr_BB93", "2dBBr002_Irreg", "2dFBr001_Irreg", "2dFBr001_TUDFoc7"]


def reference_case(name, tmp_path):
    import importlib.util
    spec = importlib.util.spec_from_file_location("mf2d", CASE2D / "make_fixtures.py")
    mf = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mf)
    return C.model_from_input(mf.materialize(name, str(tmp_path / name)))


@pytest.mark.parametrize("name", REF_CASES)
def test_reference_case_rhs_parity_full_grid(tmp_path, name):
    """Device (full-grid path) against the oracle on the case's own setup, on
    every one of the 6 nx ny values; breaking runs (tprev < t)."""
    m = reference_case(name, tmp_path)
    say("reference case rhs", name, m.nx, m.ny, "order", m.order, "flat", m.flat)
    y = W.random_state(m, 5, amp=0.3 * m.meta["hs"], kfrac=0.3)
    o = O.Oracle2D(m)
    o.set_tprev(0.0)
    ref = o.rhs(7.3, y)
    for prec, tol in ((HW_PREC_FP64, 1e-11), (HW_PREC_FP32, 3e-5)):
        dev = Hawassi2D(m, prec)
        assert dev.full_grid
        dev.set_tprev(0.0)
        e = relerr(dev.rhs(7.3, y), ref)
        dev.close()
        say("  prec", prec, "relL2", e)
        assert e < tol


@pytest.mark.parametrize("name", REF_CASES)
def test_reference_case_trajectory_full_grid(tmp_path, name):
    """From rest with the case's influx, 1.5 s (FP64): the same steps as the
    oracle and the same state to 1e-9."""
    m = reference_case(name, tmp_path)
    n = m.nx * m.ny
    y = np.zeros(6 * n)
    o = O.Oracle2D(m)
    dev = Hawassi2D(m, HW_PREC_FP64)
    dev.set_state(y)
    o.set_tprev(0.0)
    dev.set_tprev(0.0)
    rc, t_o, h_o, so = o.integrate(0.0, 1.5, 0.1, y, m.eps_abs, m.eps_rel)
    t_d, h_d = dev.integrate(0.0, 1.5, 0.1)
    st = dev.stats()
    got = dev.get_state()
    dev.close()
    e = relerr(got, y)
    say("  ", name, "steps oracle", so, "device", st["steps_accepted"], st["steps_rejected"], "relL2", e)
    assert rc == 0 and t_o == t_d == 1.5
    assert (st["steps_accepted"], st["steps_rejected"]) == tuple(so)
    assert e < 1e-9


def _case_with_walls(tmp_path, src, name, wall_line, walltype_line, bath=None):
    d = tmp_path / name
    d.mkdir()
    out = []
    for ln in (GOLD / src).read_text().split("\n"):
        key = ln.split("=")[0].strip() if "=" in ln else ""
        out.append(wall_line if key == "wall" else walltype_line if key == "walltype" else ln)
    (d / "input.txt").write_text("\n".join(out))
    if bath is not None:
        (d / "bath.dat").write_text("\n".join(" ".join("%.7e" % v for v in row) for row in bath))
    return C.model_from_input(str(d))


@pytest.mark.parametrize("which", ["flat_R05", "flat_R1", "two_depth_R05"])
def test_case_rectangle_walls_parity(tmp_path, which):
    """Walls given as input.txt rectangles (fun_wallchar's rectangle branch,
    2d/functions.c:285-345): flat bottom with R = 0.5 and R = 1 (both
    fun_delphiH1hat wall branches, on the full-grid path) and the two-depth
    fused path with R = 0.5; device against the oracle on the case's setup,
    and 1.5 s of a random sea in FP64 on the full-grid path with the oracle's
    steps."""
    if which.startswith("flat"):
        r = "0.5" if which == "flat_R05" else "1"
        m = _case_with_walls(tmp_path, "2dF001harm.txt", which, "wall \t\t= 2 600 0 660 250\n 900 300 950 400",
                             "walltype \t= 0 %s energy" % r)
    else:
        nx, ny = 512, 256
        xx = np.linspace(0, 3000, nx)
        yy = np.linspace(-250, 1500, ny)
        depth = 30.0 - 20.0 * xx[None, :] / 3000.0 + 2.0 * np.sin(yy[:, None] / 200.0)
        m = _case_with_walls(tmp_path, "2dB001HarmRefr.txt", which, "wall \t\t= 1 1500 -250 1560 900",
                             "walltype \t= 0 0.5 energy", bath=depth)
    assert m.wall
    say("rectangle walls", which, m.nx, m.ny, "flat", m.flat, "order", m.order)
    y = W.random_state(m, 5, amp=0.5, kfrac=0.3)
    ref = O.Oracle2D(m).rhs(37.3, y)
    for prec, tol in ((HW_PREC_FP64, 1e-11), (HW_PREC_FP32, 3e-5)):
        dev = Hawassi2D(m, prec)
        got = dev.rhs(37.3, y)
        if not dev.full_grid:
            got, ref_c = HS.oracle_half(got, m.nx, m.ny), HS.oracle_half(ref, m.nx, m.ny)
        else:
            ref_c = ref
        dev.close()
        e = relerr(got, ref_c)
        say("  prec", prec, "relL2", e)
        assert e < tol
    # the trajectory on the full-grid path, whose error norm covers the
    # reference's whole state (the oracle's step sequence), from a sea state
    # that already covers the walls
    y0 = W.random_state(m, 9, amp=0.1, kfrac=0.2)
    o = O.Oracle2D(m)
    dev = Hawassi2D(m, HW_PREC_FP64, full_grid=True)
    dev.set_state(y0)
    rc, t_o, h_o, so = o.integrate(0.0, 1.5, 0.1, y0, m.eps_abs, m.eps_rel)
    t_d, h_d = dev.integrate(0.0, 1.5, 0.1)
    st = dev.stats()
    got = dev.get_state()
    dev.close()
    e = relerr(got, y0)
    say("  steps oracle", so, "device", st["steps_accepted"], st["steps_rejected"], "relL2", e)
    assert rc == 0 and t_o == t_d == 1.5
    assert (st["steps_accepted"], st["steps_rejected"]) == tuple(so)
    assert e < 1e-9
