"""Input and output formats of the 2-D reference program, so that its cases
and its post-processing run unmodified around this library.

* ``read_input``: the ``input.txt`` parser (read_input, 2d/functions.c:131-223).
  The reference reads the file with a fixed sequence of ``fgets`` and
  ``fscanf`` calls; ``CStream`` reproduces the C library semantics of the
  directives it uses (``%s``, ``%d``, ``%lf``, ``%*s``, whitespace, matching
  failures that leave the offending character unread), so a file is read
  exactly as the reference reads it -- including the case where a one-word
  ``kinematic`` line shifts the next two reads and leaves ``ngauge`` at 0.
* ``read_bathy``: fun_bathy2d's data part (2d/functions.c:801-841).
* ``read_gauges``: the gauge index computation (2d/functions.c:760-784).
* ``write_constant``: ``Constant_XX.dat`` (print_xt, 2d/fun_odesolver.c:2532-2563).
* ``FieldWriter``: ``Hawassi_eta_XX.dat`` / ``_u_`` / ``_v_`` float32 Ny x Nx
  frames (2d/fun_odesolver.c:2687-2750).
* ``GaugeWriter``: the three gauge text files (print_data_gauges_uv,
  2d/fun_odesolver.c:427-476).
* ``udw2p_setup`` / ``udw2p_write``: the UDW2p header and records
  (2d/UDW2p_HAWASSI2D.c:8-123).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

_WS = b" \t\n\v\f\r"


class ScanMismatch(Exception):
    """A conversion found no matching input (C's matching failure)."""


class CStream:
    """A byte stream read with the semantics of C's ``fgets`` / ``fscanf``."""

    def __init__(self, data: bytes):
        self.b = data
        self.p = 0

    def eof(self) -> bool:
        return self.p >= len(self.b)

    def skip_ws(self):
        while self.p < len(self.b) and self.b[self.p] in _WS:
            self.p += 1

    def fgets(self, size: int = 256) -> Optional[str]:
        """Up to size-1 bytes, through the first newline."""
        if self.eof():
            return None
        end = self.b.find(b"\n", self.p)
        end = len(self.b) if end < 0 else end + 1
        end = min(end, self.p + size - 1)
        s = self.b[self.p:end]
        self.p = end
        return s.decode("latin-1")

    def token(self) -> str:
        """``%s``: skip whitespace, read non-whitespace."""
        self.skip_ws()
        if self.eof():
            raise ScanMismatch("input failure")
        q = self.p
        while q < len(self.b) and self.b[q] not in _WS:
            q += 1
        s = self.b[self.p:q].decode("latin-1")
        self.p = q
        return s

    def integer(self) -> int:
        """``%d``: optional sign and decimal digits."""
        self.skip_ws()
        q = self.p
        if q < len(self.b) and self.b[q] in b"+-":
            q += 1
        d0 = q
        while q < len(self.b) and 48 <= self.b[q] <= 57:
            q += 1
        if q == d0:
            # the sign, if any, was consumed by the conversion
            self.p = d0 if d0 > self.p else self.p
            raise ScanMismatch("integer expected")
        v = int(self.b[self.p:q])
        self.p = q
        return v

    def real(self) -> float:
        """``%lf``: the longest prefix strtod accepts (decimal, exponent,
        inf/nan)."""
        self.skip_ws()
        s = self.b[self.p:self.p + 512].decode("latin-1")
        i = 0
        if i < len(s) and s[i] in "+-":
            i += 1
        low = s[i:i + 8].lower()
        for word in ("infinity", "inf", "nan"):
            if low.startswith(word):
                self.p += i + len(word)
                return float(s[:i] + word)
        j = i
        while j < len(s) and s[j].isdigit():
            j += 1
        if j < len(s) and s[j] == ".":
            j += 1
            while j < len(s) and s[j].isdigit():
                j += 1
        if j == i or s[i:j] == ".":
            self.p += min(j, i + (1 if s[i:j] == "." else 0))
            raise ScanMismatch("real expected")
        k = j
        if k < len(s) and s[k] in "eE":
            m = k + 1
            if m < len(s) and s[m] in "+-":
                m += 1
            if m < len(s) and s[m].isdigit():
                while m < len(s) and s[m].isdigit():
                    m += 1
                k = m
        self.p += k
        return float(s[:k])

    def scan(self, fmt: str) -> list:
        """Run one fscanf format; returns the converted values (a shorter list
        after a matching failure, as fscanf stops there)."""
        out = []
        i = 0
        while i < len(fmt):
            c = fmt[i]
            if c in " \t\n":
                self.skip_ws()
                i += 1
                continue
            if c != "%":
                if self.eof() or chr(self.b[self.p]) != c:
                    return out
                self.p += 1
                i += 1
                continue
            i += 1
            suppress = fmt[i] == "*"
            if suppress:
                i += 1
            conv = fmt[i]
            if conv == "l":
                i += 1
                conv = "f"
            i += 1
            try:
                v = {"s": self.token, "d": self.integer, "f": self.real}[conv]()
            except ScanMismatch:
                return out
            if not suppress:
                out.append(v)
        return out


@dataclass
class InputParams:
    """Everything read_input fills, under the reference's names."""
    wavename: str = ""
    ramp: float = 0.0
    taper: float = 0.0
    initial: str = ""
    t_start: float = 0.0
    t_end: float = 0.0
    dt: float = 0.0
    Tp: float = 0.0
    Hs: float = 0.0
    gamJS: float = 0.0
    s: str = ""
    current: tuple = (0.0, 0.0)
    mdir: float = 0.0                 # radians (the file holds degrees)
    orientation: str = ""
    propagation: str = ""
    x_start: float = 0.0
    y_start: float = 0.0
    x_end: float = 0.0
    y_end: float = 0.0
    line_influx: tuple = (0.0, 0.0, 0.0, 0.0)   # xstart, ystart, xend, yend
    dx: float = 0.0
    dy: float = 0.0
    damp: tuple = (0.0, 0.0, 0.0, 0.0)          # Lleft, Lright, Lbottom, Ltop
    bathy_id: str = ""
    depth: float = 0.0
    Dmin: float = 0.0
    Nadj: float = 0.0
    model: str = ""
    breaking: str = ""
    friction: str = ""
    kinematic: str = ""
    kinematic_type: str = ""
    ngauge: int = 0
    npartition: int = 0
    KBC: float = 0.0
    TC: float = 0.0
    Tchar: float = 0.0
    kin_interval: tuple = (0.0, 0.0, 0.0, 0.0)  # xinterv1, yinterv1, xinterv2, yinterv2
    kin_rec: tuple = (0.0, 0.0)                 # dx_rec, dy_rec
    nsurf: int = 0
    ndeep: int = 0
    cut: str = ""
    ninterp: int = 0
    mid: str = ""
    min_z: str = ""
    max_z: str = ""
    wall: str = ""
    walls: List[tuple] = field(default_factory=list)   # (x0, y0, x1, y1)
    walltype: int = 0
    refl_coef: float = 0.0
    wallmethod: str = ""
    coupling: str = ""
    oz_start: tuple = (0.0, 0.0)
    oz_end: tuple = (0.0, 0.0)
    oz_in: tuple = (0.0, 0.0)
    oz_tr: tuple = (0.0, 0.0)
    dir_force: str = ""
    # derived exactly as read_input derives them
    delb: float = 1.5
    U: float = 0.0
    cutfrac: int = 0

    @property
    def order(self) -> int:
        """The dynamics as the library's model code (1, 2 or 3)."""
        return {"HS1": 1, "HS2": 2, "HS3": 3}.get(self.model, 0)


def _set(obj, names, vals):
    for n, v in zip(names, vals):
        if n is not None:
            setattr(obj, n, v)


def read_input(path_or_bytes, stream=CStream) -> InputParams:
    """read_input (2d/functions.c:131-223): the same reads in the same order.
    A value whose read fails keeps its initial value (0 / empty), as the
    reference's globals do.  `stream` builds the reader from the file's bytes
    (any object with CStream's fgets and scan)."""
    data = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
    cs = stream(bytes(data))
    p = InputParams()
    sc = cs.scan
    cs.fgets()
    _set(p, ["wavename"], sc("%*s %*s %s \n"))
    _set(p, ["ramp", "taper"], sc("%*s %*s %lf %lf \n"))
    _set(p, ["initial"], sc("%*s %*s %s \n"))
    cs.fgets()
    _set(p, ["t_start", "t_end", "dt"], sc("%*s %*s %lf \n %*s %*s %lf \n %*s %*s %lf\n"))
    _set(p, ["Tp", "Hs", "gamJS", "s"], sc("%*s %*s %lf \n %*s %*s %lf \n %*s %*s %lf \n %*s %*s %s\n"))
    v = sc("%*s %*s %lf %lf \n %*s %*s %lf\n")
    cur = list(p.current)
    for i, x in enumerate(v[:2]):
        cur[i] = x
    p.current = tuple(cur)
    mdir = v[2] if len(v) > 2 else 0.0
    _set(p, ["orientation", "propagation"], sc("%*s %*s %s %s\n"))
    cs.fgets()
    _set(p, ["x_start", "y_start", "x_end", "y_end"], sc("%*s %*s %lf %lf\n %*s %*s %lf %lf\n"))
    v = sc("%*s %*s %lf %lf %lf %lf\n")
    p.line_influx = tuple(v + [0.0] * (4 - len(v)))
    v = sc("%*s %*s %lf %lf\n %*s %*s %lf %lf %lf %lf\n ")
    _set(p, ["dx", "dy"], v[:2])
    if len(v) > 2:
        p.damp = tuple(v[2:] + [0.0] * (6 - len(v)))
    _set(p, ["bathy_id", "depth", "Dmin"], sc("%*s %*s %s \n %*s %*s %lf \n %*s %*s %lf \n"))
    _set(p, ["Nadj"], sc("%*s %*s %lf \n"))
    cs.fgets()
    _set(p, ["model"], sc("%*s %*s %s \n"))
    _set(p, ["breaking"], sc("%*s %*s %s \n"))
    _set(p, ["friction"], sc("%*s %*s %s \n"))
    _set(p, ["kinematic", "kinematic_type"], sc("%*s %*s %s %s \n"))
    _set(p, ["ngauge"], sc("%*s %*s %d \n"))
    _set(p, ["npartition"], sc("%*s %*s %d \n"))
    cs.fgets()
    _set(p, ["KBC", "TC", "Tchar"], sc("%*s %*s %lf \n %*s %*s %lf %lf \n"))
    cs.fgets()
    v = sc("%*s %*s %lf %lf\n %*s %*s %lf %lf\n")
    p.kin_interval = tuple(v + [0.0] * (4 - len(v)))
    v = sc("%*s %*s %lf %lf\n")
    p.kin_rec = tuple(v + [0.0] * (2 - len(v)))
    _set(p, ["nsurf", "ndeep"], sc("%*s %*s %d \n %*s %*s %d \n"))
    cs.fgets()
    _set(p, ["cut"], sc("%*s %*s %s \n"))
    _set(p, ["ninterp", "mid"], sc("%*s %*s %d %s\n"))
    _set(p, ["min_z"], sc("%*s %*s %s \n"))
    _set(p, ["max_z"], sc("%*s %*s %s \n"))
    _set(p, ["wall"], sc("%*s %*s %s "))
    # the reference's test is true for every value, so any word reads as
    # atoi(word) wall segments (0 for "no")
    nwall = _atoi(p.wall)
    for _ in range(max(nwall, 0)):
        v = sc("%lf %lf %lf %lf")
        p.walls.append(tuple(v + [0.0] * (4 - len(v))))
    sc("\n")
    _set(p, ["walltype", "refl_coef", "wallmethod"], sc("%*s %*s %d %lf %s\n"))
    cs.fgets()
    _set(p, ["coupling"], sc("%*s %*s %s \n"))
    for name in ("oz_start", "oz_end", "oz_in", "oz_tr"):
        v = sc("%*s %*s %lf %lf\n")
        setattr(p, name, tuple(v + [0.0] * (2 - len(v))))
    _set(p, ["dir_force"], sc("%*s %*s %s \n"))
    p.delb = 1.5
    p.mdir = mdir * math.pi / 180.0
    p.U = math.sqrt(p.current[0] ** 2 + p.current[1] ** 2)
    if p.cut == "default":
        p.cutfrac = {"HS1": 2, "HS2": 4, "HS3": 6}.get(p.model, 0)
    else:
        p.cutfrac = _atoi(p.cut)
    return p


def _atoi(s: str) -> int:
    """C's atoi: leading whitespace, sign, digits; 0 when none."""
    s = s.lstrip()
    i = 0
    if i < len(s) and s[i] in "+-":
        i += 1
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    return int(s[:j]) if j > i else 0


def _atof(s: str) -> float:
    """C's atof: the longest prefix strtod accepts; 0.0 when none."""
    try:
        return CStream(s.encode("latin-1")).real()
    except ScanMismatch:
        return 0.0


def c_round(v: float) -> int:
    """C's round(): halves away from zero."""
    return int(math.copysign(math.floor(abs(v) + 0.5), v))


def read_bathy(path, nx: int, ny: int, bathy_id: str, depth: float):
    """fun_bathy2d (2d/functions.c:801-841): the bathymetry field [ny][nx]
    (negative below the still water level: flat gives -depth, user_defined
    the negated values of bath.dat read row by row) and runup_id (1 when any
    point is above the water).  A short file stops at the first failed read,
    leaving the remaining points at 0 as the reference's zero-initialised
    array does."""
    b = np.zeros((ny, nx))
    if bathy_id == "flat":
        b[:] = -depth
    elif bathy_id == "user_defined":
        cs = CStream(open(path, "rb").read())
        done = False
        for i in range(ny):
            for j in range(nx):
                v = cs.scan("%lf ")
                if not v:
                    done = True
                    break
                b[i, j] = -v[0]
            if done:
                break
    runup_id = 1 if b.max() > 0 else 0
    return b, runup_id


def read_gauges(path, ngauge: int, x: np.ndarray, y: np.ndarray, dx: float, dy: float):
    """Gauge positions (x, y per line of gauges.dat) to grid indices
    round((x - x0) / dx), round((y - y0) / dy) (2d/functions.c:760-784)."""
    cs = CStream(open(path, "rb").read())
    idx = np.zeros((2, ngauge), dtype=np.int64)
    for j in range(ngauge):
        v = cs.scan("%lf %lf\n")
        xg, yg = (v + [0.0, 0.0])[:2]
        idx[0, j] = c_round((xg - x[0]) / dx)
        idx[1, j] = c_round((yg - y[0]) / dy)
        if idx[0, j] >= len(x) or idx[1, j] >= len(y):
            raise ValueError("Gauges are outside of the domain!")
    return idx


# ------------------------------------------------------------------ output --
def write_constant(path, x, y, th, hs: float, kp: float, wp: float, kx, ky, bathy, cfsa, chiadj=None):
    """Constant_XX.dat (print_xt, 2d/fun_odesolver.c:2532-2563): native-endian
    float64 records Nx, x[Nx], Ny, y[Ny], Nt, t[Nt], Hs, kp, wp, kx[Nx],
    ky[Ny], bathymetry [Ny][Nx], cfSA [Ny][Nx] and, unless the influx is
    "zero", ChiAdj [Ny][Nx]."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    th = np.asarray(th, np.float64)
    with open(path, "wb") as f:
        for a in ([float(len(x))], x, [float(len(y))], y, [float(len(th))], th, [hs, kp, wp],
                  np.asarray(kx, np.float64), np.asarray(ky, np.float64),
                  np.asarray(bathy, np.float64).reshape(len(y), len(x)),
                  np.asarray(cfsa, np.float64).reshape(len(y), len(x))):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        if chiadj is not None:
            f.write(np.ascontiguousarray(chiadj, dtype=np.float64).reshape(len(y), len(x)).tobytes())


def read_constant(path, nx_hint: Optional[int] = None, with_chi: bool = True) -> dict:
    """Reads a Constant_XX.dat back (the inverse of write_constant)."""
    a = np.fromfile(path, dtype=np.float64)
    out = {}
    i = 0
    nx = int(a[i]); i += 1
    out["x"] = a[i:i + nx]; i += nx
    ny = int(a[i]); i += 1
    out["y"] = a[i:i + ny]; i += ny
    nt = int(a[i]); i += 1
    out["t"] = a[i:i + nt]; i += nt
    out["Hs"], out["kp"], out["wp"] = a[i:i + 3]; i += 3
    out["kx"] = a[i:i + nx]; i += nx
    out["ky"] = a[i:i + ny]; i += ny
    out["bathy"] = a[i:i + nx * ny].reshape(ny, nx); i += nx * ny
    out["cfsa"] = a[i:i + nx * ny].reshape(ny, nx); i += nx * ny
    if with_chi and i < len(a):
        out["chiadj"] = a[i:i + nx * ny].reshape(ny, nx); i += nx * ny
    return out


class FieldWriter:
    """Hawassi_eta_XX.dat, Hawassi_u_XX.dat, Hawassi_v_XX.dat (2d/fun_odesolver.c:
    2687-2750): a new partition (XX = 01, 02, ...) opens every ntpart output
    times; each output time appends the surface elevation as float32 [Ny][Nx],
    and u, v only at the first output time of a partition when the kinematic
    module is off, at every output time when it is on."""

    def __init__(self, outdir: str, ntpart: int, kinematic: bool = False):
        self.outdir = outdir
        self.ntpart = ntpart
        self.kinematic = kinematic
        self.i = 0
        self.ipart = 1
        self.f = None

    def _open(self):
        self.close()
        name = lambda v: f"{self.outdir}Hawassi_{v}_{self.ipart:02d}.dat"
        self.f = [open(name(v), "wb") for v in ("eta", "u", "v")]
        self.ipart += 1

    def write(self, eta, u, v):
        first = self.i % self.ntpart == 0
        if first:
            self._open()
        self.f[0].write(np.ascontiguousarray(eta, dtype=np.float32).tobytes())
        if (not self.kinematic and first) or self.kinematic:
            self.f[1].write(np.ascontiguousarray(u, dtype=np.float32).tobytes())
            self.f[2].write(np.ascontiguousarray(v, dtype=np.float32).tobytes())
        for f in self.f:
            f.flush()
        self.i += 1

    def close(self):
        if self.f:
            for f in self.f:
                f.close()
        self.f = None


class GaugeWriter:
    """The three gauge text files of print_data_gauges_uv
    (2d/fun_odesolver.c:427-476): at the first output two header rows
    ("0 " and the gauge x, then "0 " and the gauge y), then one row per output
    time: t and the value at each gauge, every number as C's "%f "."""

    def __init__(self, files, index_gauge, x, y):
        self.f = files
        self.idx = np.asarray(index_gauge)
        self.x = x
        self.y = y
        self.i = 0

    def write(self, t: float, eta, u, v):
        jj, ii = self.idx[0], self.idx[1]
        if self.i == 0:
            hx = "0 " + "".join("%f " % self.x[j] for j in jj) + "\n"
            hy = "0 " + "".join("%f " % self.y[i] for i in ii) + "\n"
            for f in self.f:
                f.write(hx)
                f.write(hy)
        for f, a in zip(self.f, (eta, u, v)):
            f.write("%f " % t + "".join("%f " % a[i, j] for i, j in zip(ii, jj)) + "\n")
            f.flush()
        self.i += 1


def _c_e(v: float, width: int, prec: int) -> str:
    """C's "%-W.PE"."""
    return ("%-*.*E" % (width, prec, v))


def _udw2p_header(filename, nheader, nx, ny, np_, nt, dx, dy, dt, xscale, tscale, depth, hs, wp, flat):
    tp = 2 * math.pi / wp
    lines = [
        "######################################\n",
        "#                                     \n",
        "#  HAWASSI 2D output (no calibration)\n",
        "#  Case Hs=%f, Tp= %.4f   \n" % (hs, tp),
        "#                                     \n",
        "######################################\n",
        "PRECISION = single\n",
        "Software = HAWASSI\n",
        "Nheader = %d\n" % nheader,
        "Nx = %d\n" % nx,
        "Ny = %d\n" % ny,
        "Np = %d\n" % np_,
        "Nt = %d\n" % nt,
        "dx = %s\n" % _c_e(dx / xscale, 16, 12),
        "dy = %s\n" % _c_e(dy / xscale, 16, 12),
        "dt = %s\n" % _c_e(dt / tscale, 16, 12),
        "xScale = %-16.12f\n" % xscale,
        "tScale = %-16.12f\n" % tscale,
        "WD = %-16.12f\n" % xscale,
    ]
    text = "".join(lines)
    if not flat:
        text += "Bathymetry = variable\n"
        d = np.asarray(depth, np.float64).reshape(ny, nx) / xscale
        text += "".join("%-10.8f " % v for v in d.ravel())
    with open(filename, "w") as f:
        f.write(text)
    return len(text.encode("latin-1"))


def udw2p_setup(filename, nx, ny, np_, nt, dx, dy, dt, xscale, tscale, depth, hs, wp, flat):
    """UDW2p_setUp_hw2D (2d/UDW2p_HAWASSI2D.c:56-80): the text header padded
    with spaces to a whole number of records; a header longer than one record
    is rewritten with Nheader = size / record + 1.  The record size is the
    reference's formula 4 (2 Nx + 2 Np) Ny + 4 (2 Ny + 2 Np) Nx, which counts
    the Ny x Nx planes twice: the header takes that many bytes per header
    record, while udw2p_write appends 4 (2 Nx Ny + 2 Ny Np + 2 Nx Np) bytes per
    time.  Returns (record bytes, Nheader)."""
    rec = (4 * (nx * 2 + np_ * 2) * ny) + (4 * (ny * 2 + np_ * 2) * nx)
    sz = _udw2p_header(filename, 1, nx, ny, np_, nt, dx, dy, dt, xscale, tscale, depth, hs, wp, flat)
    nheader = sz // rec + 1
    if nheader > 1:
        sz = _udw2p_header(filename, nheader, nx, ny, np_, nt, dx, dy, dt, xscale, tscale, depth, hs, wp, flat)
    with open(filename, "ab") as f:
        if sz < rec * nheader:
            f.write(b" " * (rec * nheader - sz))
    return rec, nheader


def udw2p_write(filename, elev, fphi, left, right, front, back):
    """UDW2p_write_hw2D (2d/UDW2p_HAWASSI2D.c:83-123): one record of float32
    elevation [Ny][Nx], surface potential [Ny][Nx], left and right boundary
    potentials [Ny][Np], front and back [Nx][Np]."""
    with open(filename, "ab") as f:
        for a in (elev, fphi, left, right, front, back):
            f.write(np.ascontiguousarray(a, dtype=np.float32).tobytes())
