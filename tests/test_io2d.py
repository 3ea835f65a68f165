"""Input and output formats of the 2-D reference program (hawassi_c_amd.io2d).

The input parser is pinned against the C library itself: the reference's
input.txt files (tests/golden/input2d/, copied from its test cases) are read
once through io2d's CStream and once through a reader whose fgets / fscanf
are glibc's own (ctypes, fmemopen), and every field must agree.  The writers
are pinned by their byte layouts and by C's printf formatting (snprintf)."""
import ctypes
import ctypes.util
import math
import os
from pathlib import Path

import numpy as np
import pytest

from hawassi_c_amd import io2d

GOLD = Path(__file__).parent / "golden" / "input2d"
CASES = sorted(p.stem for p in GOLD.glob("*.txt"))

libc = ctypes.CDLL(ctypes.util.find_library("c"))
libc.fmemopen.restype = ctypes.c_void_p
libc.fmemopen.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p]
libc.fgets.restype = ctypes.c_char_p
libc.fgets.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p]
libc.fclose.argtypes = [ctypes.c_void_p]
libc.feof.argtypes = [ctypes.c_void_p]


class GlibcStream:
    """CStream's interface with glibc's fgets and fscanf underneath."""

    def __init__(self, data: bytes):
        self.buf = ctypes.create_string_buffer(data, len(data))
        self.fp = libc.fmemopen(self.buf, len(data), b"r")
        assert self.fp

    def __del__(self):
        if getattr(self, "fp", None):
            libc.fclose(self.fp)

    def fgets(self, size=256):
        b = ctypes.create_string_buffer(size)
        r = libc.fgets(b, size, self.fp)
        return None if r is None else b.value.decode("latin-1")

    def scan(self, fmt):
        conv = []
        i = 0
        while i < len(fmt):
            if fmt[i] == "%":
                i += 1
                sup = fmt[i] == "*"
                if sup:
                    i += 1
                c = fmt[i]
                if c == "l":
                    i += 1
                    c = "f"
                if not sup:
                    conv.append(c)
            i += 1
        objs = [ctypes.create_string_buffer(512) if c == "s" else ctypes.c_int() if c == "d" else ctypes.c_double()
                for c in conv]
        args = [o if isinstance(o, ctypes.Array) else ctypes.byref(o) for o in objs]
        n = libc.fscanf(ctypes.c_void_p(self.fp), fmt.encode(), *args)
        out = []
        for o in objs[:max(n, 0)]:
            out.append(o.value.decode("latin-1") if isinstance(o, ctypes.Array) else o.value)
        return out


def fields(p):
    return {k: v for k, v in vars(p).items()}


@pytest.mark.parametrize("case", CASES)
def test_read_input_matches_glibc(case):
    data = (GOLD / f"{case}.txt").read_bytes()
    a = io2d.read_input(data)
    b = io2d.read_input(data, stream=GlibcStream)
    fa, fb = fields(a), fields(b)
    for k in fa:
        assert fa[k] == fb[k], (case, k, fa[k], fb[k])


def kv(text):
    """An independent line reader: first token -> the tokens after '='."""
    out = {}
    for line in text.splitlines():
        t = line.split()
        if len(t) >= 3 and t[1] == "=":
            out.setdefault(t[0], t[2:])
    return out


@pytest.mark.parametrize("case", CASES)
def test_read_input_values(case):
    text = (GOLD / f"{case}.txt").read_text()
    d = kv(text)
    p = io2d.read_input((GOLD / f"{case}.txt"))
    assert p.wavename == d["Influx_signal"][0]
    assert p.model == d["dynamics"][0] and p.order in (1, 2, 3)
    assert p.breaking == d["breaking"][0]
    assert p.bathy_id == d["bathymetry"][0]
    assert (p.dx, p.dy) == tuple(float(v) for v in d["dx"][:2])
    assert p.damp == tuple(float(v) for v in d["Damp"][:4])
    assert p.line_influx == tuple(float(v) for v in d["Line_influx"][:4])
    assert (p.x_start, p.y_start) == tuple(float(v) for v in d["Xstart"][:2])
    assert (p.t_start, p.t_end, p.dt) == (float(d["T0"][0]), float(d["Tend"][0]), float(d["dt"][0]))
    assert (p.Tp, p.Hs) == (float(d["Tp"][0]), float(d["Hs"][0]))
    assert p.ninterp == int(d["depth_ref"][0]) and p.mid == d["depth_ref"][1]
    assert p.cutfrac == {"HS1": 2, "HS2": 4, "HS3": 6}[p.model]
    assert p.npartition == int(d["npartition"][0])
    assert p.mdir == float(d["direction"][0]) * math.pi / 180
    assert p.KBC == float(d["initiation"][0])
    assert (p.TC, p.Tchar) == tuple(float(v) for v in d["termination"][:2])
    assert p.wall == "no" and p.walls == []
    assert (p.walltype, p.refl_coef, p.wallmethod) == (int(d["walltype"][0]), float(d["walltype"][1]), d["walltype"][2])
    # a one-word kinematic line takes the next word as kinematic_type and
    # shifts the ngauge read onto the npartition line (ngauge stays 0)
    if len(d["kinematic"]) == 1:
        assert p.kinematic_type == "ngauge" and p.ngauge == 0
    else:
        assert p.kinematic_type == d["kinematic"][1] and p.ngauge == int(d["ngauge"][0])


def test_read_input_shift_and_walls_match_glibc():
    """Two wall segments, a gauge count hidden by the one-word kinematic line,
    a non-default cutfrac and a mid depth given as a number."""
    text = (GOLD / "2dBBr001_HarmBr_BB93.txt").read_text()
    text = text.replace("ngauge \t\t= 0", "ngauge \t\t= 3").replace("cutfrac\t\t= default", "cutfrac\t\t= 5")
    text = text.replace("wall \t\t= no", "wall \t\t= 2 0 0 10 0\n-1.5 2 3e1 4.25")
    assert "= 2 0 0 10 0" in text and "= 5" in text and "= 3" in text
    a = io2d.read_input(text.encode())
    b = io2d.read_input(text.encode(), stream=GlibcStream)
    assert fields(a) == fields(b)
    assert a.ngauge == 0 and a.cutfrac == 5 and a.walls == [(0.0, 0.0, 10.0, 0.0), (-1.5, 2.0, 30.0, 4.25)]
    assert a.ninterp == 3 and a.mid == "0.25"


@pytest.mark.parametrize("s", [b"  -12abc", b"+", b"-.5e3x", b".x", b"1e5", b"7 8", b"nan", b"-inf 3", b"x1"])
def test_scan_primitives_match_glibc(s):
    for fmt in ("%d%s", "%lf%s", "%s %lf"):
        a = io2d.CStream(s)
        b = GlibcStream(s)
        va, vb = a.scan(fmt), b.scan(fmt)
        assert len(va) == len(vb), (s, fmt, va, vb)
        for x, y in zip(va, vb):
            assert (x == y) or (isinstance(x, float) and math.isnan(x) and math.isnan(y)), (s, fmt, va, vb)


def test_read_bathy_and_gauges(tmp_path):
    nx, ny = 5, 3
    vals = np.arange(1, nx * ny + 1, dtype=float) * 0.5
    (tmp_path / "bath.dat").write_text(" ".join("%.7e" % v for v in vals) + "\n")
    b, runup = io2d.read_bathy(tmp_path / "bath.dat", nx, ny, "user_defined", 0.0)
    assert np.array_equal(b, -vals.reshape(ny, nx)) and runup == 0
    b, runup = io2d.read_bathy(None, nx, ny, "flat", 7.0)
    assert np.all(b == -7.0) and runup == 0
    (tmp_path / "bath.dat").write_text("1 2 -3")     # short file, a point above water
    b, runup = io2d.read_bathy(tmp_path / "bath.dat", nx, ny, "user_defined", 0.0)
    assert b.ravel()[:3].tolist() == [-1.0, -2.0, 3.0] and np.all(b.ravel()[3:] == 0) and runup == 1
    x = np.linspace(0, 40, 5)
    y = np.linspace(-10, 10, 3)
    (tmp_path / "gauges.dat").write_text("15 0\n 5 -10\n25.0001 9.9\n")
    idx = io2d.read_gauges(tmp_path / "gauges.dat", 3, x, y, 10.0, 10.0)
    # 1.5 rounds away from zero (C round), not to even
    assert idx.tolist() == [[2, 1, 3], [1, 0, 2]]
    (tmp_path / "gauges.dat").write_text("45 0\n")
    with pytest.raises(ValueError):
        io2d.read_gauges(tmp_path / "gauges.dat", 1, x, y, 10.0, 10.0)


def test_constant_file_layout(tmp_path):
    nx, ny, nt = 6, 4, 3
    x, y, th = np.arange(nx) * 2.0, np.arange(ny) * 3.0 - 1, np.arange(nt) * 0.5
    kx, ky = np.arange(nx) * 0.1, np.arange(ny) * 0.2
    bathy, cfsa, chi = (np.random.default_rng(s).standard_normal((ny, nx)) for s in (1, 2, 3))
    f = tmp_path / "Constant_01.dat"
    io2d.write_constant(f, x, y, th, 1.5, 0.02, 0.45, kx, ky, bathy, cfsa, chi)
    raw = np.fromfile(f, dtype=np.float64)
    assert raw.size == 3 + 2 * (nx + ny) + nt + 3 + 3 * nx * ny
    assert raw[0] == nx and raw[1 + nx] == ny and raw[2 + nx + ny] == nt
    r = io2d.read_constant(f)
    assert np.array_equal(r["x"], x) and np.array_equal(r["y"], y) and np.array_equal(r["t"], th)
    assert (r["Hs"], r["kp"], r["wp"]) == (1.5, 0.02, 0.45)
    for k, a in (("kx", kx), ("ky", ky), ("bathy", bathy), ("cfsa", cfsa), ("chiadj", chi)):
        assert np.array_equal(r[k], a)
    io2d.write_constant(f, x, y, th, 1.5, 0.02, 0.45, kx, ky, bathy, cfsa, None)
    assert np.fromfile(f, dtype=np.float64).size == 3 + 2 * (nx + ny) + nt + 3 + 2 * nx * ny


def test_field_writer_partitions(tmp_path):
    nx, ny = 4, 3
    w = io2d.FieldWriter(str(tmp_path) + "/", ntpart=3, kinematic=False)
    frames = [np.full((ny, nx), i, dtype=float) for i in range(7)]
    for fr in frames:
        w.write(fr, fr + 100, fr + 200)
    w.close()
    fsz = 4 * nx * ny
    sizes = {p.name: p.stat().st_size for p in tmp_path.iterdir()}
    assert sizes == {"Hawassi_eta_01.dat": 3 * fsz, "Hawassi_eta_02.dat": 3 * fsz, "Hawassi_eta_03.dat": fsz,
                     "Hawassi_u_01.dat": fsz, "Hawassi_u_02.dat": fsz, "Hawassi_u_03.dat": fsz,
                     "Hawassi_v_01.dat": fsz, "Hawassi_v_02.dat": fsz, "Hawassi_v_03.dat": fsz}
    e2 = np.fromfile(tmp_path / "Hawassi_eta_02.dat", dtype=np.float32).reshape(3, ny, nx)
    assert [int(e2[i, 0, 0]) for i in range(3)] == [3, 4, 5]
    assert np.fromfile(tmp_path / "Hawassi_u_02.dat", dtype=np.float32)[0] == 103


def c_printf(fmt, v):
    b = ctypes.create_string_buffer(128)
    libc.snprintf(b, 128, fmt.encode(), ctypes.c_double(v))
    return b.value.decode()


def test_gauge_writer_text(tmp_path):
    x = np.array([0.0, 1.25, 2.5, 3.75])
    y = np.array([-1.0, 0.5, 2.0])
    idx = np.array([[1, 3], [0, 2]])
    fs = [open(tmp_path / f"g{i}.txt", "w") for i in range(3)]
    g = io2d.GaugeWriter(fs, idx, x, y)
    eta = np.arange(12.0).reshape(3, 4) / 7
    g.write(0.1, eta, -eta, 2 * eta)
    g.write(1e-7, eta, -eta, 2 * eta)
    for f in fs:
        f.close()
    lines = (tmp_path / "g1.txt").read_text().splitlines(keepends=True)
    assert lines[0] == "0 " + c_printf("%f ", 1.25) + c_printf("%f ", 3.75) + "\n"
    assert lines[1] == "0 " + c_printf("%f ", -1.0) + c_printf("%f ", 2.0) + "\n"
    assert lines[2] == c_printf("%f ", 0.1) + c_printf("%f ", -eta[0, 1]) + c_printf("%f ", -eta[2, 3]) + "\n"
    assert lines[3].startswith(c_printf("%f ", 1e-7))


def test_udw2p_layout(tmp_path):
    nx, ny, npz, nt = 8, 4, 3, 5
    f = str(tmp_path / "out.udw2p")
    depth = np.full((ny, nx), 12.5)
    rec, nh = io2d.udw2p_setup(f, nx, ny, npz, nt, 2.0, 3.0, 0.1, 12.5, 1.0, depth, 1.0, 0.5, flat=False)
    assert rec == 4 * (2 * nx + 2 * npz) * ny + 4 * (2 * ny + 2 * npz) * nx
    assert os.path.getsize(f) == rec * nh and nh >= 1
    text = open(f).read()
    assert "#  Case Hs=%s, Tp= %s   \n" % (c_printf("%f", 1.0), c_printf("%.4f", 2 * math.pi / 0.5)) in text
    assert "dx = %s\n" % c_printf("%-16.12E", 2.0 / 12.5) in text
    assert "xScale = %s\n" % c_printf("%-16.12f", 12.5) in text
    assert "Nheader = %d\n" % nh in text and "Bathymetry = variable\n" in text
    assert text.count(c_printf("%-10.8f ", 1.0)) == nx * ny
    io2d.udw2p_write(f, np.zeros((ny, nx)), np.ones((ny, nx)), np.zeros((ny, npz)), np.zeros((ny, npz)),
                     np.zeros((nx, npz)), np.full((nx, npz), 2.0))
    # the header is sized by the reference's record-size formula, which counts
    # the two Ny x Nx planes twice; a written record holds the planes once
    data = 4 * (2 * nx * ny + 2 * ny * npz + 2 * nx * npz)
    assert data < rec
    assert os.path.getsize(f) == rec * nh + data
    last = np.fromfile(f, dtype=np.float32, offset=rec * nh)
    assert last.size * 4 == data and last[-1] == 2.0 and last[nx * ny] == 1.0
    # a small record forces a multi-record header with Nheader rewritten
    f2 = str(tmp_path / "small.udw2p")
    rec2, nh2 = io2d.udw2p_setup(f2, 2, 2, 1, 1, 1.0, 1.0, 1.0, 1.0, 1.0, np.ones((2, 2)), 1.0, 1.0, flat=False)
    assert nh2 > 1 and os.path.getsize(f2) == rec2 * nh2 and ("Nheader = %d\n" % nh2) in open(f2).read()
