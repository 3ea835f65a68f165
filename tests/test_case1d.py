"""The 1-D reference case path (hawassi_c_amd.io1d, hawassi_c_amd.case1d).

* The input.txt reads are pinned against glibc's own fgets / fscanf on all
  eighteen of the reference's 1-D case files (tests/golden/input1d/).
* The case set-up of 1dB002HarmBar is pinned by the reference's own output
  of that case: the processed influx.dat its program wrote (ramped user
  record, ramp length from the detected peak period) must come out byte for
  byte, and the peak frequency, wave number and wave length its logfile
  reports must agree.
* The Area generation kernel is checked against the closed form of the DFT
  of a rolled sequence; the gauge spline against scipy's natural cubic
  spline; harmonic_influx's sample count rule on small cases.
"""
import gzip
import math
import os
import shutil
from pathlib import Path

import numpy as np
import pytest
import scipy.interpolate as si

from hawassi_c_amd import case1d, io1d
from hawassi_c_amd import setup1d as S
from test_io2d import GlibcStream

GOLD = Path(__file__).parent / "golden"
CASES = sorted(p.stem for p in (GOLD / "input1d").glob("*.txt"))


def unpack_case(name: str, dst: Path) -> Path:
    """The fixture case with its .gz files expanded, as the reference reads it."""
    src = GOLD / "case1d" / name
    out = dst / name
    for root, _, files in os.walk(src):
        for f in files:
            rel = Path(root).relative_to(src)
            (out / rel).mkdir(parents=True, exist_ok=True)
            if f.endswith(".gz"):
                (out / rel / f[:-3]).write_bytes(gzip.decompress((Path(root) / f).read_bytes()))
            else:
                shutil.copyfile(Path(root) / f, out / rel / f)
    return out


def test_eighteen_cases_present():
    assert len(CASES) == 18


@pytest.mark.parametrize("case", CASES)
def test_read_input_matches_glibc(case):
    data = (GOLD / "input1d" / f"{case}.txt").read_bytes()
    a = io1d.read_input(data)
    b = io1d.read_input(data, stream=GlibcStream)
    assert vars(a) == vars(b)
    assert a.dynmodel in ("HS1", "HS2", "HS3") and a.propagation == "Uni+"
    assert a.dx > 0 and a.Xend > a.Xstart


def test_read_input_fields():
    p = io1d.read_input(GOLD / "input1d" / "1dB002HarmBar.txt")
    assert (p.wavename, p.initial, p.bath, p.dynmodel, p.influxing) == ("user_defined", "zero", "user_defined",
                                                                      "HS2", "Area")
    assert (p.Xstart, p.Xend, p.Xinflux, p.dampL, p.dampR) == (-20.0, 50.0, 5.7, 20.0, 20.0)
    assert p.dx == 0.03419638495359062 and p.adjcoef == 0.3 and p.ngauge == 6
    assert (p.rampbool, p.nTramp, p.cutfracwn) == ("yes", 8.0, 4)
    # user_defined records: the time window is not read
    assert (p.T0, p.Tend, p.dt) == (0.0, 0.0, 0.0)
    # "ramp_influx = no" has no count: the %lf read fails on the next line's
    # first word, which the wall read then takes
    q = io1d.read_input(GOLD / "input1d" / "1dFBr005Bor.txt")
    assert (q.rampbool, q.nTramp, q.wall, q.walltype) == ("no", 0.0, "no", 1)
    assert (q.wavename, q.initial, q.Tend, q.dt, q.cutfracwn) == ("zero", "user_defined", 120.0, 0.02, 6)
    r = io1d.read_input(GOLD / "input1d" / "1dRBr001Harm.txt")
    assert r.mid_temp == "6.666666666666667" and r.cutfracwn == 8


@pytest.mark.parametrize("name", ["1dB002HarmBar", "1dB001IrSlope"])
def test_processed_influx_is_the_references(tmp_path, name):
    """The processed influx.dat (ramped record, ramp length from the detected
    peak period) byte for byte as the reference's program wrote it: a
    harmonic record (5000 samples, ramp 8 periods) and an irregular one
    (91969 samples, ramp 4 periods)."""
    d = unpack_case(name, tmp_path)
    c = case1d.case_from_dir(str(d))
    assert c.model.n == 2048                           # logfiles: Number of Nodes
    out = tmp_path / "influx.dat"
    io1d.write_influx(out, c.t_out, c.insig)
    assert out.read_bytes() == (d / "expected" / "influx.dat").read_bytes()


def test_b002_derived_values(tmp_path):
    d = unpack_case("1dB002HarmBar", tmp_path)
    c = case1d.case_from_dir(str(d))
    # logfile: peak frequency 3.10 rad/s, kp 1.68, wave length 3.75 m,
    # steepness kp Hs/2 0.02
    dv = c.derived
    assert round(dv["nu_peak"], 2) == 3.10 and round(dv["kp"], 2) == 1.68
    assert round(dv["lambda_p"], 2) == 3.75 and round(dv["kp"] * dv["Hs_in"] / 2, 2) == 0.02
    # three reference depths for the varying bottom
    assert c.model.ops["gam_mid"] is not None and not c.model.flat


def test_area_kernel_closed_form():
    """x_space's Area kernel is the DFT of Re(IFFT(Ug)) rolled by s nodes:
    F(roll(Re F^-1 Ug, s))_m = e^{-2 pi i m s / N} (Ug_m + conj Ug_{-m}) / 2,
    normalised so that its m = 0 entry is sqrt(g D)/dx."""
    n, L, D = 256, 40.0, 0.5
    x, dx = S.domain_n(-10.0, -10.0 + L, n)
    k = S.freqspace1d(L, n)
    xin = float(x[S.closest(x, 3.3)])
    g = case1d.area_gamx(k, x, xin, x[0], dx, D, 2.0, short=False)
    ug = np.array([S.group_velocity_i(float(v), D) for v in k])
    s = int((xin - x[0]) / dx)
    m = np.arange(n)
    want = np.exp(-2j * np.pi * m * s / n) * (ug + np.conj(ug[(-m) % n])) / 2
    want *= math.sqrt(S.GRAV * D) / dx / want[0].real
    assert np.allclose(g, want, rtol=1e-12, atol=1e-12 * abs(want[0]))
    assert abs(g[0] - math.sqrt(S.GRAV * D) / dx) < 1e-12 * abs(g[0])


def test_gauge_spline_is_natural_cubic():
    x = np.linspace(-20, 50, 700)
    y = np.sin(0.7 * x) + 0.2 * np.cos(2.9 * x)
    g = np.array([-20.0, 10.5, 12.5, 17.3, 49.99, 50.0])
    w = io1d.gauge_weights(x, g)
    assert np.abs(w @ y - si.CubicSpline(x, y, bc_type="natural")(g)).max() < 1e-13


def test_gauge_file_layout(tmp_path):
    x = np.linspace(0, 10, 11)
    gw = io1d.GaugeWriter1D(tmp_path / "g.txt", np.array([2.5, 7.0]), x)
    gw.write(0.0, np.zeros(11))
    gw.write(0.03931, x * 0.001)
    gw.close()
    txt = (tmp_path / "g.txt").read_text().splitlines()
    assert txt[0] == "0 2.500000 7.000000 "
    assert txt[1] == "0.000000 0.000000 0.000000 "
    assert txt[2] == "0.039310 0.002500 0.007000 "
    head, t, v = io1d.read_gauge_file(tmp_path / "g.txt")
    assert list(head) == [2.5, 7.0] and v.shape == (2, 2)


def test_harmonic_record_sample_rule():
    # samples continue while the previous one is <= Tend: 0, .5, 1, 1.5
    t, e = case1d.harmonic_record(0.0, 1.0, 2.0, 0.5, 0.3)
    assert list(t) == [0.0, 0.5, 1.0, 1.5]
    assert abs(e[1] - 0.3 * (1 - math.cos(0.5 * math.pi / 8)) / 2 * math.sin(0.5 * math.pi)) < 1e-15


def test_ramp_edges():
    t = np.arange(0, 40.0, 0.25)
    r = case1d.influx_ramp(t, 2.0, 4.0)
    assert r[0] == 0.0 and r[-1] == 0.0
    assert np.all(r[(t > 8.0) & (t < t[-1] - 8.0)] == 1.0)
    assert abs(r[8] - (1 - math.cos(2.0 * math.pi / 8.0)) / 2) < 1e-15
    with pytest.raises(ValueError):
        case1d.influx_ramp(t, 0.0, 4.0)


def test_unsupported_options_raise(tmp_path):
    # friction (1dRBr001Harm) and runup (1dR001Harm: a bathymetry above the
    # still water level) are refused
    for case in ("1dRBr001Harm", "1dR001Harm"):
        d = tmp_path / case
        d.mkdir()
        shutil.copyfile(GOLD / "input1d" / f"{case}.txt", d / "input.txt")
        (d / "bath.dat").write_text("".join("%g\n" % v for v in np.linspace(5.0, -5.0, 5000)))
        with pytest.raises(NotImplementedError):
            case1d.case_from_dir(str(d))


def _oracle_gauge_rows(case_dir: Path, nsamples: int):
    """The FP64 oracle run through the first nsamples output times of a case,
    the gauge values printed as the reference prints them ("%f")."""
    from oracle import oracle as O
    c = case1d.case_from_dir(str(case_dir))
    o = O.Oracle1D(c.model)
    y = c.y0.copy()
    t, h = float(c.t_out[0]), 1e-4
    n = c.model.n
    w = io1d.gauge_weights(c.x, c.gauges)
    rows = []
    for i, ti in enumerate(c.t_out[:nsamples]):
        if i > 0:
            rc, t, h, _ = o.integrate(t, float(ti), h, y)
            assert rc == 0
        eta = np.real(np.fft.ifft(y[:n] + 1j * y[n:2 * n]))
        rows.append("%f " % ti + "".join("%f " % v for v in w @ eta))
    return rows


@pytest.mark.parametrize("name,nsamples", [("1dB002HarmBar", 800), ("1dFBr005Bor", 400)])
def test_oracle_reproduces_reference_record(tmp_path, name, nsamples):
    """The oracle from the case's own input files reproduces the gauge record
    the reference's program wrote, character for character, over the first
    nsamples output times (31 s of the bar case with the waves through all six
    gauges; 8 s of the bore case, with breaking, through its gauge)."""
    d = unpack_case(name, tmp_path)
    ref = (d / "expected" / "Hawassi_eta_gauge.txt").read_text().splitlines()
    got = _oracle_gauge_rows(d, nsamples)
    assert max(abs(float(v)) for r in ref[1:nsamples + 1] for v in r.split()[1:]) > 0.01
    assert got == [r for r in ref[1:nsamples + 1]]


def _wall_case(tmp_path, name="1dF005WallFreq", refl=None):
    """A wall case from the reference's input file with a synthetic record
    (the reference ships none for these cases): a seeded ramped sum of
    sines sampled every 0.1 s."""
    d = tmp_path / name
    d.mkdir(exist_ok=True)
    txt = (GOLD / "input1d" / f"{name}.txt").read_text()
    if refl is not None:
        txt = txt.replace("walltype 	= 2 2500 1", f"walltype 	= 2 2500 {refl}")
    (d / "input.txt").write_text(txt)
    t = np.arange(0, 400.0, 0.1)
    rs = np.random.default_rng(5)
    e = sum(0.02 * rs.uniform(0.5, 1) * np.sin(2 * math.pi * t / tp + rs.uniform(0, 6.28)) for tp in (10, 12, 15))
    (d / "influx.dat").write_text("".join("%.6f %.8f\n" % (a, b) for a, b in zip(t, e)))
    (d / "gauges.dat").write_text("100\n")
    return case1d.case_from_dir(str(d))


def test_wall_kernel_and_chi():
    """define_wall's kernel is the Area kernel rolled to the wall node,
    times i omega(k), negated for a wall left of the influx; wall_setup
    multiplies ChiAdj by 1 - cf(Xwall - d) + cf(Xwall)."""
    n, L, D = 256, 40.0, 0.5
    x, dx = S.domain_n(-10.0, -10.0 + L, n)
    k = S.freqspace1d(L, n)
    for xwall, sign in ((20.0, 1.0), (-5.0, -1.0)):
        p = io1d.Input1D(Xwall=xwall, refl_coef=0.7, adjcoef=0.5)
        m = S.Model1D(n=n, x=x, dx=dx, k=k, aal=np.ones(n), chi=np.ones(n), damp=np.zeros(n), depth=np.full(n, D),
                      flat=True)
        w = case1d.wall_setup(p, m, x, k, dx, D, 1.7, 2.0, 0.0)
        idx = S.closest(x, xwall)
        ug = np.array([S.group_velocity_i(float(v), D) for v in k])
        mm = np.arange(n)
        g = np.exp(-2j * np.pi * mm * idx / n) * (ug + np.conj(ug[(-mm) % n])) / 2
        g *= math.sqrt(S.GRAV * D) / dx / g[0].real
        want = sign * 1j * S.omega_v(k, D) * g
        assert w["idx"] == idx and w["refl"] == 0.7
        assert np.allclose(w["gamw_skew"], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        ramp = 1 - S.cf(x, xwall - 1.0, 1.0) + S.cf(x, xwall, 1.0)
        assert np.array_equal(m.chi, ramp)
        assert m.chi[x < xwall - 1.0].min() == 1.0 and m.chi[x > xwall + 1.0].min() == 1.0
        assert m.chi[idx] < 1.0


def test_oracle_wall_source_history(tmp_path):
    """fun_wallinfluxsource in the oracle over a scripted call sequence: the
    eta equation gains aal * 2 S gamW_skew with S the trapezoid integral of
    refl * eta(x_wall) over the times of the calls that advanced time (the
    first such call sets S = 0; calls at an earlier or equal time reuse S);
    the u equation is unchanged; refl = 0 gives no source."""
    from oracle import oracle as O
    c = _wall_case(tmp_path)
    m = c.model
    w = m.wall
    no = S.Model1D(**{**m.__dict__, "wall": None})
    ow, on = O.Oracle1D(m), O.Oracle1D(no)
    n = m.n
    rs = np.random.default_rng(11)
    calls = [0.5, 0.9, 0.9, 0.7, 1.4, 2.0]
    S_, tprev, prev, it = 0.0, float(c.t_out[0]), 0.0, 0
    for t in calls:
        eta = 0.01 * rs.standard_normal(n)
        eh = np.fft.fft(eta)
        y = np.concatenate([eh.real, eh.imag, 0.1 * eh.real, 0.1 * eh.imag])
        if t > tprev:
            se = w["refl"] * np.real(np.fft.ifft(eh))[w["idx"]]
            S_ = 0.0 if it < 1 else S_ + (se + prev) * (t - tprev) / 2
            it, tprev, prev = it + 1, t, se
        d = ow.rhs(t, y) - on.rhs(t, y)
        want = m.aal * 2 * S_ * w["gamw_skew"]
        scale = np.abs(ow.rhs(t, y)[:2 * n]).max()
        assert np.abs(d[:n] - want.real).max() <= 1e-12 * scale
        assert np.abs(d[n:2 * n] - want.imag).max() <= 1e-12 * scale
        assert np.abs(d[2 * n:]).max() == 0.0
    assert S_ != 0.0
    c0 = _wall_case(tmp_path / "r0", refl=0) if (tmp_path / "r0").mkdir() is None else None
    o0 = O.Oracle1D(c0.model)
    on0 = O.Oracle1D(S.Model1D(**{**c0.model.__dict__, "wall": None}))
    for t in calls:
        assert np.array_equal(o0.rhs(t, y), on0.rhs(t, y))


def _oracle_break_rows(case_dir: Path, t_end: float):
    """breaking.dat rows of the FP64 oracle over [t0, t_end], printed as
    data_saving_break prints them ("%lf %d " and "%d " per node)."""
    from oracle import oracle as O
    c = case1d.case_from_dir(str(case_dir))
    o = O.Oracle1D(c.model, rec_dt=float(c.t_out[1] - c.t_out[0]))
    y = c.y0.copy()
    t, h = float(c.t_out[0]), 1e-4
    for ti in c.t_out[1:]:
        if ti > t_end + 1e-12:
            break
        rc, t, h, _ = o.integrate(t, float(ti), h, y)
        assert rc == 0
    return ["%lf %d " % (tt, len(nd)) + "".join("%d " % v for v in nd) for tt, nd in o.break_records()]


@pytest.mark.parametrize("name,t_end", [("1dFBr005Bor", 13.0), ("1dBBr002IrBar", 16.0)])
def test_oracle_reproduces_reference_breaking_records(tmp_path, name, t_end):
    """The breaking.dat the reference's program wrote for two breaking cases
    (the bore from initial.dat over a flat bottom; irregular waves breaking on
    a bar, AreaShort generation, three reference depths), row for row: the
    time of the RHS call that logs (an RKF45 stage time, six decimals) and
    every node the breaking loop marks.  It pins the breaking criterion, the
    crest tracking and the stepper's stage times together."""
    d = unpack_case(name, tmp_path)
    ref = [r for r in (d / "expected" / "breaking.dat").read_text().splitlines() if float(r.split()[0]) <= t_end]
    got = _oracle_break_rows(d, t_end)
    assert len(ref) >= 20
    assert got == ref


@pytest.mark.parametrize("name", ["1dB002HarmBar", "1dFBr005Bor", "1dB001IrSlope", "1dBBr002IrBar"])
def test_constants_file_is_the_references(tmp_path, name):
    """Constants_01.dat (print_xt: Nx, x, Nt, the output times, depth, k,
    nu_peak as raw doubles) byte for byte as the reference's program wrote it
    for these cases: the grid, the record times, bath.dat, freqspace's k and
    the peak frequency (from the influx record's spectrum; for the bore, from
    the initial elevation's)."""
    d = unpack_case(name, tmp_path)
    c = case1d.case_from_dir(str(d))
    m = c.model
    out = tmp_path / "out"
    fw = io1d.FieldWriter1D(str(out), c.x, c.t_out, m.depth, m.k, c.derived["nu_peak"], c.params.npartition)
    n = m.n
    fw.write(np.zeros(n), np.zeros(n), np.zeros(n))
    fw.close()
    assert (out / "Constants_01.dat").read_bytes() == (d / "expected" / "Constants_01.dat").read_bytes()
    assert (out / "Hawassi_eta_01.dat").stat().st_size == 8 * n


def test_field_writer_partitions(tmp_path):
    """n = 11 output times in npartition = 3 parts: 3 + 3 + 3 and a remainder
    part of 2 (Constants_04 holds the remainder times)."""
    x = np.linspace(0, 1, 8)
    t = np.arange(11) * 0.5
    fw = io1d.FieldWriter1D(str(tmp_path), x, t, np.ones(8), np.arange(8.0), 2.5, 3)
    for i in range(11):
        fw.write(np.full(8, i), np.full(8, -i), np.full(8, 2 * i))
    fw.close()
    for ip, tp in ((1, t[0:3]), (2, t[3:6]), (3, t[6:9]), (4, t[9:])):
        c = np.fromfile(tmp_path / f"Constants_{ip:02d}.dat", "<f8")
        assert c[0] == 8 and np.array_equal(c[1:9], x) and c[9] == len(tp)
        assert np.array_equal(c[10:10 + len(tp)], tp) and c[-1] == 2.5
        e = np.fromfile(tmp_path / f"Hawassi_eta_{ip:02d}.dat", "<f8").reshape(-1, 8)
        assert np.array_equal(e[:, 0], tp / 0.5)
    assert not (tmp_path / "Constants_05.dat").exists()
