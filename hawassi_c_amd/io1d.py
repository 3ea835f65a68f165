"""Input and output formats of the 1-D reference program, so that its cases
run unmodified around this library.

* ``read_input``: the ``input.txt`` reads of read_input_file
  (1d/read_input.c:1-56) followed by those of influxlib
  (1d/influx_lib.c:1-107).  Both functions skip to a fixed line with
  ``fgets`` (after a ``rewind``) and then ``fscanf``; the reads are replayed
  here in the same order with ``CStream``, which reproduces C's ``fgets`` and
  ``fscanf`` semantics (a failed conversion leaves its variable at the
  initial value and the offending word unread, e.g. ``ramp_influx = no``
  with no period count).
* ``read_signal``: read_data's two-column influx.dat (1d/functions.c:932-957).
* ``read_column``: bath.dat and gauges.dat (1d/read_input.c:91-106,
  1d/functions.c:832-852).
* ``read_initial``: initial.dat, Nx rows of eta and u (1d/read_input.c:62-75).
* ``write_influx``: the processed influx ``influx.dat`` in the output folder
  (read_influx_file, 1d/read_input.c:81-89), "%g %g" per sample.
* ``write_constant`` / ``FieldWriter1D``: ``Constants_XX.dat`` and the
  full-domain ``Hawassi_eta/u/phi_XX.dat`` frames, per output partition
  (print_xt, 1d/functions.c:604-666, 784-802).
* ``GaugeWriter1D``: ``Hawassi_eta_gauge.txt`` (ode_solver_Ham's gauge block
  and print_data_gauge, 1d/functions.c:500-526, 700-744): a header row
  "0 " plus each gauge position, then per output time t and the natural
  cubic spline of eta(x) at each gauge, every number as C's "%f ".
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from .io2d import CStream, _atoi


@dataclass
class Input1D:
    """The globals read_input_file and influxlib fill, under the reference's
    names (1d/header.h).  Values whose read fails keep these initial
    values, as the reference's zero-initialised globals do."""
    wavename: str = ""
    initial: str = ""
    T0: float = 0.0
    Tend: float = 0.0
    dt: float = 0.0
    Tp: float = 0.0
    Hs: float = 0.0
    gamJS: float = 0.0
    Xstart: float = 0.0
    Xend: float = 0.0
    Xinflux: float = 0.0
    dx: float = 0.0
    dampL: float = 0.0
    dampR: float = 0.0
    bath: str = ""
    depthflat: float = 0.0
    adjcoef: float = 0.0
    dynmodel: str = ""
    breaking: str = ""
    friction: str = ""
    influxing: str = ""
    propagation: str = ""
    kinematic: str = ""
    ngauge: int = 0
    npartition: int = 0
    parKBC: float = 0.0
    parTC: float = 0.0
    parTchar: float = 0.0
    xinterv1: float = 0.0
    xinterv2: float = 0.0
    nsurf: int = 0
    ndeep: int = 0
    cut_temp: str = ""
    mid_temp: str = ""
    min_z: str = ""
    max_z: str = ""
    rampbool: str = ""
    nTramp: float = 0.0
    wall: str = ""
    walltype: int = 0
    Xwall: float = 0.0
    refl_coef: float = 0.0
    coupling: str = ""
    dir_force: str = ""
    x_cpl: tuple = (0.0, 0.0)
    alpha_corr: float = 0.0

    @property
    def cutfracwn(self) -> int:
        """influxlib's cutfrac by dynamics (1d/influx_lib.c:79-90), replaced by
        the expert option when it is not "default" (read_kinematics,
        1d/functions.c:3556-3558, which runs before alias)."""
        if self.cut_temp != "default":
            return _atoi(self.cut_temp)
        return {"HS1": 2, "HS2": 4, "HS3": 6, "HS4": 8}.get(self.dynmodel, 0)

    @property
    def order(self) -> int:
        """The dynamics as the library's model code (1, 2 or 3)."""
        return {"HS1": 1, "HS2": 2, "HS3": 3}.get(self.dynmodel, 0)


def _set(obj, names, vals):
    for n, v in zip(names, vals):
        setattr(obj, n, v)


def _skip(cs, nlines: int, size: int):
    for _ in range(nlines):
        cs.fgets(size)


def read_input(path_or_bytes, stream=CStream) -> Input1D:
    """read_input_file (1d/read_input.c:1-56), then the input.txt reads of
    influxlib (1d/influx_lib.c:1-76), in the reference's order."""
    data = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
    data = bytes(data)
    p = Input1D()
    # read_input_file: fgets buffers of 128 bytes
    cs = stream(data)
    cs.fgets(128)
    _set(p, ["wavename"], cs.scan("%*s %*s %s \n"))
    _set(p, ["initial"], cs.scan("%*s %*s %s \n"))
    cs = stream(data)
    _skip(cs, 18, 128)
    _set(p, ["bath", "depthflat", "adjcoef"], cs.scan("%*s %*s %s \n %*s %*s %lf \n %*s %*s %lf \n"))
    cs = stream(data)
    _skip(cs, 23, 128)
    for name, fmt in (("dynmodel", "%*s %*s %s \n"), ("breaking", "%*s %*s %s \n"),
                      ("friction", "%*s %*s %s \n"), ("influxing", "%*s %*s %s \n"),
                      ("propagation", "%*s %*s %s \n"), ("kinematic", "%*s %*s %s \n"),
                      ("ngauge", "%*s %*s %d \n"), ("npartition", "%*s %*s %d \n")):
        _set(p, [name], cs.scan(fmt))
    cs.fgets(128)
    _set(p, ["parKBC", "parTC", "parTchar"], cs.scan("%*s %*s %lf \n %*s %*s %lf %lf \n"))
    cs.fgets(128)
    _set(p, ["xinterv1", "xinterv2"], cs.scan("%*s %*s %lf \n %*s %*s %lf \n"))
    _set(p, ["nsurf", "ndeep"], cs.scan("%*s %*s %d \n %*s %*s %d \n"))
    cs.fgets(128)
    _set(p, ["cut_temp"], cs.scan("%*s %*s %s \n"))
    _set(p, ["mid_temp"], cs.scan("%*s %*s %s \n"))
    _set(p, ["min_z"], cs.scan("%*s %*s %s \n"))
    _set(p, ["max_z"], cs.scan("%*s %*s %s \n"))
    _set(p, ["rampbool", "nTramp"], cs.scan("%*s %*s %s %lf"))
    _set(p, ["wall"], cs.scan("%*s %*s %s \n"))
    _set(p, ["walltype", "Xwall", "refl_coef"], cs.scan("%*s %*s %d %lf %lf\n"))
    cs.fgets(128)
    v = cs.scan("%*s %*s %s %s\n %*s %*s %lf %lf\n")
    _set(p, ["coupling", "dir_force"], v[:2])
    if len(v) > 2:
        p.x_cpl = tuple(v[2:] + [0.0] * (4 - len(v)))
    _set(p, ["alpha_corr"], cs.scan("%*s %*s %lf\n"))
    # influxlib: fgets buffers of 256 bytes; the time window is read only
    # for the generated signals, and the file is rewound after it
    if p.wavename in ("harmonic", "jonswap", "zero"):
        cs = stream(data)
        _skip(cs, 5, 256)
        _set(p, ["T0", "Tend", "dt"], cs.scan("%*s %*s %lf \n %*s %*s %lf \n %*s %*s %lf\n"))
        _set(p, ["Tp", "Hs", "gamJS"], cs.scan("%*s %*s %lf \n %*s %*s %lf \n %*s %*s %lf\n"))
    cs = stream(data)
    _skip(cs, 13, 256)
    _set(p, ["Xstart", "Xend", "Xinflux"], cs.scan("%*s %*s %lf \n %*s %*s %lf \n %*s %*s %lf\n"))
    _set(p, ["dx", "dampL", "dampR"], cs.scan("%*s %*s %lf \n %*s %*s %lf %lf\n "))
    if p.rampbool == "no":
        p.nTramp = 0.0                        # 1d/main.cc:47-49
    return p


def read_signal(path) -> tuple[np.ndarray, np.ndarray]:
    """read_data (1d/functions.c:932-957): pairs "%lf %lf" until the read
    reaches the end of the file."""
    cs = CStream(open(path, "rb").read())
    t, s = [], []
    while True:
        v = cs.scan("%lf %lf")
        if len(v) < 2:
            break
        t.append(v[0])
        s.append(v[1])
    return np.asarray(t), np.asarray(s)


def read_column(path, n: int) -> np.ndarray:
    """n values read as "%lf \\n" (bath.dat, gauges.dat).  A short file leaves
    the remaining entries at 0."""
    cs = CStream(open(path, "rb").read())
    out = np.zeros(n)
    for j in range(n):
        v = cs.scan("%lf \n")
        if not v:
            break
        out[j] = v[0]
    return out


def read_initial(path, n: int) -> tuple[np.ndarray, np.ndarray]:
    """read_initial_file (1d/read_input.c:62-75): n rows "%lf %lf\\n"."""
    cs = CStream(open(path, "rb").read())
    eta, u = np.zeros(n), np.zeros(n)
    for j in range(n):
        v = cs.scan("%lf %lf\n")
        if len(v) > 0:
            eta[j] = v[0]
        if len(v) > 1:
            u[j] = v[1]
        if len(v) < 2:
            break
    return eta, u


def c_g(v: float) -> str:
    """C's "%g"."""
    return "%g" % v


def write_influx(path, t: np.ndarray, sig: np.ndarray):
    """read_influx_file (1d/read_input.c:81-89)."""
    with open(path, "w") as f:
        for a, b in zip(t, sig):
            f.write("%s %s\n" % (c_g(a), c_g(b)))


def gauge_weights(x: np.ndarray, gauges: np.ndarray) -> np.ndarray:
    """The natural cubic spline through (x, y) evaluated at the gauges, as a
    matrix W with W @ y = s(gauges) (the spline is linear in y): GSL's
    cspline, which print_data_gauge uses (1d/functions.c:500-526)."""
    from .setup2d import spline_coeffs
    n = len(x)
    eye = np.eye(n)
    c = spline_coeffs(x, eye)
    W = np.empty((len(gauges), n))
    for gi, t in enumerate(np.asarray(gauges, dtype=np.float64)):
        i = int(np.clip(np.searchsorted(x, t, side="right") - 1, 0, n - 2))
        h = x[i + 1] - x[i]
        d = t - x[i]
        b = (eye[:, i + 1] - eye[:, i]) / h - h * (c[:, i + 1] + 2.0 * c[:, i]) / 3.0
        dd = (c[:, i + 1] - c[:, i]) / (3.0 * h)
        W[gi] = eye[:, i] + d * (b + d * (c[:, i] + d * dd))
    return W


class GaugeWriter1D:
    """Hawassi_eta_gauge.txt (1d/functions.c:700-744, 500-526)."""

    def __init__(self, path, gauges: np.ndarray, x: np.ndarray):
        self.f = open(path, "w")
        self.g = np.asarray(gauges, dtype=np.float64)
        self.W = gauge_weights(x, self.g)
        self.f.write("0 " + "".join("%f " % v for v in self.g) + "\n")

    def write(self, t: float, eta: np.ndarray) -> np.ndarray:
        v = self.W @ eta
        self.f.write("%f " % t + "".join("%f " % a for a in v) + "\n")
        return v

    def close(self):
        if self.f:
            self.f.close()
            self.f = None


def write_constant(path, x: np.ndarray, t_part: np.ndarray, depth: np.ndarray, k: np.ndarray, nu_peak: float):
    """Constants_XX.dat (print_xt, 1d/functions.c:784-802): doubles Nx, x[Nx],
    Nt, t[Nt] (the partition's output times), depth[Nx], k[Nx], nu_peak."""
    with open(path, "wb") as f:
        for a in ([float(len(x))], x, [float(len(t_part))], t_part, depth, k, [nu_peak]):
            np.asarray(a, dtype="<f8").tofile(f)


class FieldWriter1D:
    """The full-domain outputs of ode_solver_Ham (1d/functions.c:604-666):
    the record's n output times are cut into npartition parts of
    NtoutIP = n / npartition times and a remainder part of Ntrest
    (1d/functions.c:3585-3598); each part has Constants_XX.dat and the frames
    Hawassi_eta_XX.dat, Hawassi_u_XX.dat, Hawassi_phi_XX.dat (Nx doubles per
    output time, no header), XX = 01, 02, ..."""

    def __init__(self, outdir, x, t_out, depth, k, nu_peak: float, npartition: int):
        import os
        self.dir = outdir
        self.x, self.t, self.depth, self.k, self.nu = x, np.asarray(t_out), depth, k, nu_peak
        n = len(self.t)
        npart = max(int(npartition), 1)
        self.ntout = n // npart
        self.ntrest = n - npart * self.ntout
        self.npart = npart
        self.ip = 0
        self.i = 0
        self.f = None
        os.makedirs(outdir, exist_ok=True)

    def _open(self):
        import os
        self.close()
        self.ip += 1
        tag = "%02d.dat" % self.ip
        n1 = self.ntout if self.ip <= self.npart else self.ntrest
        # print_xt writes trest whenever N1 == Ntrest (1d/functions.c:792-797)
        if n1 == self.ntrest:
            tp = self.t[self.npart * self.ntout:]
        else:
            tp = self.t[(self.ip - 1) * self.ntout:self.ip * self.ntout]
        write_constant(os.path.join(self.dir, "Constants_" + tag), self.x, tp, self.depth, self.k, self.nu)
        self.f = [open(os.path.join(self.dir, f"Hawassi_{v}_" + tag), "wb") for v in ("eta", "u", "phi")]

    def write(self, eta, u, phi):
        if self.ntout > 0 and self.i % self.ntout == 0:
            self._open()
        for f, a in zip(self.f, (eta, u, phi)):
            np.asarray(a, dtype="<f8").tofile(f)
        self.i += 1

    def close(self):
        if self.f:
            for f in self.f:
                f.close()
            self.f = None


class KinWriter1D:
    """The kinematics outputs of ode_solver_Ham (1d/functions.c:655-698,
    save_* at 1d/functions.c:400-500, kinematic_modul's fwrites at
    1d/kinematic_modul.c:228-246): per partition XX (the same cut as
    FieldWriter1D) Hawassi_Elev_XX.dat (header N1*N2, N1, N2, times, xout;
    then eta on the window per output time), Hawassi_Pressure_XX.dat,
    Hawassi_Vel_XX.dat and Hawassi_Acc_XX.dat (header N1*N2*N3, N1, N2, N3,
    times, xout, zout; then per output time P [nx][nz], dxPhi then dzPhi,
    d/dt dxPhi then d/dt dzPhi), all little-endian doubles."""

    def __init__(self, outdir, t_out, npartition: int, xout: np.ndarray, zout: np.ndarray):
        self.dir = outdir
        self.t = np.asarray(t_out)
        n = len(self.t)
        self.npart = max(int(npartition), 1)
        self.ntout = n // self.npart
        self.ntrest = n - self.npart * self.ntout
        self.xout, self.zout = np.asarray(xout, "<f8"), np.asarray(zout, "<f8")
        self.ip = 0
        self.i = 0
        self.f = None

    def _open(self):
        import os
        self.close()
        self.ip += 1
        tag = "%02d.dat" % self.ip
        n1 = self.ntout if self.ip <= self.npart else self.ntrest
        if n1 == self.ntrest:
            tp = self.t[self.npart * self.ntout:]
        else:
            tp = self.t[(self.ip - 1) * self.ntout:self.ip * self.ntout]
        n2, n3 = float(self.xout.size), float(self.zout.size)
        self.f = {}
        for name in ("Elev", "Pressure", "Vel", "Acc"):
            f = open(os.path.join(self.dir, f"Hawassi_{name}_" + tag), "wb")
            if name == "Elev":
                head = [n1 * n2, n1, n2]
            else:
                head = [n1 * n2 * n3, n1, n2, n3]
            np.asarray(head, "<f8").tofile(f)
            np.asarray(tp, "<f8").tofile(f)
            self.xout.tofile(f)
            if name != "Elev":
                self.zout.tofile(f)
            self.f[name] = f

    def write(self, eta_cut, kin: dict):
        if self.ntout > 0 and self.i % self.ntout == 0:
            self._open()
        w = lambda name, a: np.asarray(a, "<f8").tofile(self.f[name])  # noqa: E731
        w("Elev", eta_cut)
        w("Pressure", kin["P"])
        w("Vel", kin["dxPhi"])
        w("Vel", kin["dzPhi"])
        w("Acc", kin["acc_x"])
        w("Acc", kin["acc_z"])
        self.i += 1

    def close(self):
        if self.f:
            for f in self.f.values():
                f.close()
            self.f = None


def read_gauge_file(path) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """A Hawassi_*_gauge.txt back as (gauge positions, t, values[t, gauge])."""
    rows = [ln.split() for ln in open(path).read().splitlines() if ln.strip()]
    head = np.asarray([float(v) for v in rows[0][1:]])
    body = np.asarray([[float(v) for v in r] for r in rows[1:]])
    return head, body[:, 0], body[:, 1:]


def case_file(case_dir, name: str) -> Optional[str]:
    import os
    p = os.path.join(case_dir, name)
    return p if os.path.exists(p) else None
