"""TEST INFRASTRUCTURE ONLY -- ctypes wrapper of the FP64 CPU oracle.

The oracle (hawassi_oracle.c, hawassi_oracle1d.c) is a plain-C restatement
of the reference's 2-D AB2 and 1-D HS2 right-hand sides and GSL-style RKF45
drivers.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg may use it, and only
as the checker; the product library never calls it.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_dp = ct.POINTER(ct.c_double)


class OrcParams(ct.Structure):
    _fields_ = [
        ("Nx", ct.c_int), ("Ny", ct.c_int), ("flat", ct.c_int), ("wall", ct.c_int), ("refl_one", ct.c_int),
        ("Cpeak", ct.c_double),
        ("kx", _dp), ("ky", _dp), ("aal", _dp), ("fbdy", _dp), ("wallchar", _dp), ("chi", _dp),
        ("C2d", _dp), ("L2d", _dp), ("gwx", _dp), ("gwy", _dp),
        ("gam_m", _dp), ("gam_p", _dp), ("csq_m", _dp), ("csq_p", _dp), ("om2_m", _dp), ("om2_p", _dp),
        ("source_on", ct.c_int), ("vertical", ct.c_int), ("propagation", ct.c_int),
        ("newGam", _dp), ("SkewGam", _dp),
        ("nline", ct.c_int), ("nt", ct.c_int),
        ("tsig", _dp), ("sig", _dp), ("skew", _dp),
        ("breaking_on", ct.c_int),
        ("KBC", ct.c_double), ("TC", ct.c_double), ("Tchar", ct.c_double), ("delb", ct.c_double),
        ("MaxEtaInit", ct.c_double), ("Tstar", ct.c_double), ("Rsearch", ct.c_double), ("U", ct.c_double),
        ("pbchar", _dp), ("depth", _dp), ("x", _dp), ("y", _dp),
        ("order", ct.c_int),
        ("gam_c", _dp), ("csq_c", _dp), ("om2_c", _dp),
    ]


class Orc1DParams(ct.Structure):
    _fields_ = [
        ("N", ct.c_int), ("flat", ct.c_int), ("source_on", ct.c_int), ("propagation", ct.c_int),
        ("breaking_on", ct.c_int),
        ("k", _dp), ("aal", _dp), ("chi", _dp), ("damp", _dp), ("csqr", _dp), ("Lop", _dp),
        ("csq_min", _dp), ("csq_plus", _dp), ("gam_min", _dp), ("gam_plus", _dp),
        ("gamx", _dp), ("gamx_skew", _dp),
        ("nt", ct.c_int), ("t_sig", _dp), ("sig", _dp), ("sig_skew", _dp),
        ("KBC", ct.c_double), ("TC", ct.c_double), ("Tstar", ct.c_double), ("MaxEtaInit", ct.c_double),
        ("delb", ct.c_double), ("dx", ct.c_double),
        ("xbstart", ct.c_int), ("xbend", ct.c_int), ("xbendL", ct.c_int), ("xbstartR", ct.c_int),
        ("lenIF", ct.c_int), ("ILL", ct.c_int), ("IFF", ct.c_int),
        ("depth", _dp), ("x", _dp),
        ("csq_mid", _dp), ("gam_mid", _dp),
        ("order", ct.c_int),
        ("wall_on", ct.c_int), ("idx_wall", ct.c_int), ("refl_coef", ct.c_double), ("t_wall0", ct.c_double),
        ("gamw_skew", _dp),
        ("rec_dt", ct.c_double),
    ]


def build():
    r = subprocess.run(["make", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stderr)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ct.CDLL(str(LIB))
        L.orc_create.argtypes = [ct.POINTER(OrcParams)]
        L.orc_create.restype = ct.c_void_p
        L.orc_destroy.argtypes = [ct.c_void_p]
        L.orc_rhs.argtypes = [ct.c_void_p, ct.c_double, _dp, _dp]
        L.orc_rhs_count.argtypes = [ct.c_void_p]
        L.orc_rhs_count.restype = ct.c_long
        L.orc_evolve_apply.argtypes = [ct.c_void_p, _dp, ct.c_double, _dp, _dp, ct.c_double, ct.c_double,
                                       ct.POINTER(ct.c_long)]
        L.orc_integrate.argtypes = L.orc_evolve_apply.argtypes
        L.orc_hermitian_project.argtypes = [_dp, ct.c_int, ct.c_int]
        L.orc_fft2.argtypes = [ct.c_int, ct.c_int, _dp, ct.c_int]
        L.orc_ifft_c2r.argtypes = [ct.c_int, ct.c_int, _dp, _dp, _dp]
        L.orc_spline_coeffs.argtypes = [_dp, _dp, ct.c_int, _dp]
        L.orc_spline_index.argtypes = [_dp, ct.c_int, ct.c_double]
        L.orc_set_tprev.argtypes = [ct.c_void_p, ct.c_double]
        L.orc_get_tprev.argtypes = [ct.c_void_p]
        L.orc_get_tprev.restype = ct.c_double
        L.orc_break_calls.argtypes = [ct.c_void_p]
        L.orc_break_calls.restype = ct.c_long
        L.orc_flagbr.argtypes = [ct.c_void_p]
        L.orc_get_B.argtypes = [ct.c_void_p, _dp]
        L.orc_num_threads.restype = ct.c_int
        L.orc1d_create.argtypes = [ct.POINTER(Orc1DParams)]
        L.orc1d_create.restype = ct.c_void_p
        L.orc1d_destroy.argtypes = [ct.c_void_p]
        L.orc1d_rhs.argtypes = [ct.c_void_p, ct.c_double, _dp, _dp]
        L.orc1d_evolve_apply.argtypes = L.orc_evolve_apply.argtypes
        L.orc1d_integrate.argtypes = L.orc_evolve_apply.argtypes
        L.orc1d_rhs_count.argtypes = [ct.c_void_p]
        L.orc1d_rhs_count.restype = ct.c_long
        L.orc1d_break_records.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_long, ct.c_long]
        L.orc1d_break_records.restype = ct.c_long
        L.orc1d_flag.argtypes = [ct.c_void_p]
        L.orc1d_ncrest.argtypes = [ct.c_void_p]
        L.orc1d_get_B.argtypes = [ct.c_void_p, _dp]
        L.orc1d_fft.argtypes = [ct.c_int, _dp, ct.c_int]
        L.orc_set_threads.argtypes = [ct.c_int]
        L.orc_set_plant.argtypes = [ct.c_int]
        L.orc1d_set_fp32_emulation.argtypes = [ct.c_int]
        L.orc_hss.argtypes = [ct.c_void_p, _dp, ct.c_int, _dp]
        L.orc_break_diag.argtypes = [ct.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_find_peaks.argtypes = [_dp, ct.c_int, ct.c_int, ct.c_double, ct.c_void_p, ct.c_void_p]
        L.orc_find_peaks.restype = ct.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Oracle2D:
    """FP64 CPU restatement of rhs_AB2_2DF_uv / rhs_AB2_2DB_uv + rkf45."""

    def __init__(self, model):
        L = lib()
        keep = {}

        def arr(name, a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep[name] = a
            return _p(a)

        p = OrcParams()
        p.Nx, p.Ny, p.flat, p.wall, p.refl_one = model.nx, model.ny, int(model.flat), int(model.wall), int(model.refl_one)
        p.order = int(getattr(model, "order", 2))
        p.Cpeak = model.cpeak
        p.kx, p.ky = arr("kx", model.kx), arr("ky", model.ky)
        p.aal, p.fbdy = arr("aal", model.aal), arr("fbdy", model.fbdy)
        p.wallchar, p.chi = arr("w", model.wallchar), arr("chi", model.chi)
        if model.flat:
            p.C2d, p.L2d = arr("c2d", model.c2d), arr("l2d", model.l2d)
            p.gwx, p.gwy = arr("gwx", model.grad_wall_x), arr("gwy", model.grad_wall_y)
        else:
            o = model.ops
            p.gam_m, p.gam_p = arr("gm", o["gam_min"]), arr("gp", o["gam_plus"])
            p.csq_m, p.csq_p = arr("cm", o["csq_min"]), arr("cp", o["csq_plus"])
            p.om2_m, p.om2_p = arr("om", o["om2_min"]), arr("op", o["om2_plus"])
            if o.get("gam_mid") is not None:
                p.gam_c, p.csq_c, p.om2_c = arr("gc", o["gam_mid"]), arr("cc", o["csq_mid"]), arr("oc", o["om2_mid"])
        p.source_on = int(model.source_on)
        if model.source_on:
            p.vertical, p.propagation = int(model.vertical), int(model.propagation)
            p.newGam, p.SkewGam = arr("ng", model.new_gam), arr("sg", model.skew_gam)
            p.nline, p.nt = model.sig.shape
            p.tsig, p.sig, p.skew = arr("ts", model.t_sig), arr("sig", model.sig), arr("sk", model.skew_sig)
        br = getattr(model, "breaking", None)
        if br is not None:
            p.breaking_on = 1
            p.KBC, p.TC, p.Tchar, p.delb = br["KBC"], br["TC"], br["Tchar"], br["delb"]
            p.MaxEtaInit, p.Tstar, p.Rsearch, p.U = br["MaxEtaInit"], br["Tstar"], br["Rsearch"], 0.0
            p.pbchar = arr("pbchar", br["char"])
            p.depth = arr("depth", model.depth)
            p.x, p.y = arr("x", model.x), arr("y", model.y)
        self._keep = keep
        self._p = p
        self.ctx = L.orc_create(ct.byref(p))
        self.model = model
        self.dim = 6 * model.nx * model.ny

    def __del__(self):
        try:
            lib().orc_destroy(self.ctx)
        except Exception:
            pass

    def rhs(self, t, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.empty_like(y)
        lib().orc_rhs(self.ctx, float(t), _p(y), _p(out))
        return out

    def evolve_apply(self, t, t1, h, y, eps_abs, eps_rel):
        tt, hh = ct.c_double(t), ct.c_double(h)
        st = (ct.c_long * 2)()
        rc = lib().orc_evolve_apply(self.ctx, ct.byref(tt), float(t1), ct.byref(hh), _p(y), eps_abs, eps_rel, st)
        return rc, tt.value, hh.value, (st[0], st[1])

    def integrate(self, t, t1, h, y, eps_abs, eps_rel):
        tt, hh = ct.c_double(t), ct.c_double(h)
        st = (ct.c_long * 2)()
        rc = lib().orc_integrate(self.ctx, ct.byref(tt), float(t1), ct.byref(hh), _p(y), eps_abs, eps_rel, st)
        return rc, tt.value, hh.value, (st[0], st[1])

    def rhs_count(self):
        return lib().orc_rhs_count(self.ctx)

    def set_tprev(self, t):
        lib().orc_set_tprev(self.ctx, float(t))

    def tprev(self):
        return lib().orc_get_tprev(self.ctx)

    def break_calls(self):
        return lib().orc_break_calls(self.ctx)

    def flagbr(self):
        return lib().orc_flagbr(self.ctx)

    def B(self):
        out = np.empty(self.model.nx * self.model.ny)
        lib().orc_get_B(self.ctx, _p(out))
        return out.reshape(self.model.ny, self.model.nx)

    def hss(self, phi, which="csq"):
        """HSSgen_2depth (or _3depth) of the real field phi [ny, nx] with the
        C^2 ("csq") or omega^2 ("om2") multipliers: full-grid spectrum."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        out = np.empty(2 * phi.size)
        lib().orc_hss(self.ctx, _p(phi), 0 if which == "csq" else 1, _p(out))
        return (out[0::2] + 1j * out[1::2]).reshape(phi.shape)

    def break_diag(self, y):
        """(eta, Hilbert eta, local wavenumber K) over the grid, as
        breaking_criterion forms them from the state y."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        sh = (self.model.ny, self.model.nx)
        e, h, k = np.empty(sh), np.empty(sh), np.empty(sh)
        lib().orc_break_diag(self.ctx, _p(y), _p(e), _p(h), _p(k))
        return e, h, k


def _cplx(a):
    a = np.asarray(a, dtype=complex)
    buf = np.empty(2 * a.size)
    buf[0::2] = a.real
    buf[1::2] = a.imag
    return buf


class Oracle1D:
    """FP64 CPU restatement of rhs_HsF1 / rhs_HsF2 / rhs_HsF3 (flat, two or three
    reference depths) + breaking + the wall source + rkf45.  rec_dt > 0 keeps
    the breaking.dat records (one per interval rec_dt, 1d/breaking.c:426-444)."""

    def __init__(self, m, rec_dt: float = 0.0):
        L = lib()
        keep = {}

        def arr(name, a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep[name] = a
            return _p(a)

        p = Orc1DParams()
        p.N, p.flat, p.source_on = m.n, int(m.flat), int(m.source_on)
        p.order = int(getattr(m, "order", 2))
        p.propagation = {"Uni+": 0, "Uni-": 1, "Bi": 2, "None": 3}[m.propagation]
        p.k, p.aal, p.chi, p.damp = arr("k", m.k), arr("aal", m.aal), arr("chi", m.chi), arr("damp", m.damp)
        if m.flat:
            p.csqr, p.Lop = arr("csqr", m.csqr), arr("lop", m.lop)
        else:
            o = m.ops
            p.csq_min, p.csq_plus = arr("cm", o["csq_min"]), arr("cp", o["csq_plus"])
            p.gam_min, p.gam_plus = arr("gm", o["gam_min"]), arr("gp", o["gam_plus"])
            if o.get("gam_mid") is not None:
                p.csq_mid, p.gam_mid = arr("cc", o["csq_mid"]), arr("gc", o["gam_mid"])
        if m.source_on:
            p.gamx, p.gamx_skew = arr("gx", _cplx(m.gamx)), arr("gs", _cplx(m.gamx_skew))
            p.nt = len(m.t_sig)
            p.t_sig, p.sig, p.sig_skew = arr("ts", m.t_sig), arr("sig", m.sig), arr("sk", m.sig_skew)
        br = m.breaking
        p.dx = m.dx
        p.depth, p.x = arr("depth", m.depth), arr("x", m.x)
        if br is not None:
            p.breaking_on = 1
            p.KBC, p.TC, p.Tstar, p.MaxEtaInit, p.delb = br["KBC"], br["TC"], br["Tstar"], br["MaxEtaInit"], br["delb"]
            for f in ("xbstart", "xbend", "xbendL", "xbstartR", "lenIF", "ILL", "IFF"):
                setattr(p, f, int(br[f]))
        w = getattr(m, "wall", None)
        if w is not None:
            p.wall_on, p.idx_wall, p.refl_coef, p.t_wall0 = 1, int(w["idx"]), float(w["refl"]), float(w["t0"])
            p.gamw_skew = arr("gw", _cplx(w["gamw_skew"]))
        p.rec_dt = float(rec_dt)
        self._keep, self._p = keep, p
        self.ctx = L.orc1d_create(ct.byref(p))
        self.model = m
        self.dim = 4 * m.n

    def __del__(self):
        try:
            lib().orc1d_destroy(self.ctx)
        except Exception:
            pass

    def rhs(self, t, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.empty_like(y)
        lib().orc1d_rhs(self.ctx, float(t), _p(y), _p(out))
        return out

    def evolve_apply(self, t, t1, h, y, eps_abs=None, eps_rel=None):
        m = self.model
        tt, hh = ct.c_double(t), ct.c_double(h)
        st = (ct.c_long * 2)()
        rc = lib().orc1d_evolve_apply(self.ctx, ct.byref(tt), float(t1), ct.byref(hh), _p(y),
                                      m.eps_abs if eps_abs is None else eps_abs,
                                      m.eps_rel if eps_rel is None else eps_rel, st)
        return rc, tt.value, hh.value, (st[0], st[1])

    def integrate(self, t, t1, h, y, eps_abs=None, eps_rel=None):
        m = self.model
        tt, hh = ct.c_double(t), ct.c_double(h)
        st = (ct.c_long * 2)()
        rc = lib().orc1d_integrate(self.ctx, ct.byref(tt), float(t1), ct.byref(hh), _p(y),
                                   m.eps_abs if eps_abs is None else eps_abs,
                                   m.eps_rel if eps_rel is None else eps_rel, st)
        return rc, tt.value, hh.value, (st[0], st[1])

    def rhs_count(self):
        return lib().orc1d_rhs_count(self.ctx)

    def break_records(self):
        """[(t, nodes)] as data_saving_break writes them."""
        L = lib()
        n = L.orc1d_break_records(self.ctx, None, None, None, 0, 0)
        t = np.zeros(max(n, 1))
        cnt = np.zeros(max(n, 1), dtype=np.int32)
        L.orc1d_break_records(self.ctx, _p(t), cnt.ctypes.data, None, n, 0)
        tot = int(cnt[:n].sum())
        nodes = np.zeros(max(tot, 1), dtype=np.int32)
        L.orc1d_break_records(self.ctx, None, None, nodes.ctypes.data, 0, tot)
        out, o = [], 0
        for i in range(n):
            out.append((float(t[i]), nodes[o:o + cnt[i]].copy()))
            o += cnt[i]
        return out

    def flag(self):
        return lib().orc1d_flag(self.ctx)

    def ncrest(self):
        return lib().orc1d_ncrest(self.ctx)

    def B(self):
        out = np.empty(self.model.n)
        lib().orc1d_get_B(self.ctx, _p(out))
        return out


def fft1d(a: np.ndarray, sign: int) -> np.ndarray:
    buf = _cplx(a)
    lib().orc1d_fft(len(a), _p(buf), sign)
    return buf[0::2] + 1j * buf[1::2]


def hermitian_project(a: np.ndarray) -> np.ndarray:
    """fftwcomplex_2d_sym on a copy of a complex [ny, nx] array."""
    ny, nx = a.shape
    buf = np.empty(2 * a.size)
    buf[0::2] = a.real.ravel()
    buf[1::2] = a.imag.ravel()
    lib().orc_hermitian_project(_p(buf), nx, ny)
    return (buf[0::2] + 1j * buf[1::2]).reshape(ny, nx)


def fft2(a: np.ndarray, sign: int) -> np.ndarray:
    ny, nx = a.shape
    buf = np.empty(2 * a.size)
    buf[0::2] = a.real.ravel()
    buf[1::2] = a.imag.ravel()
    lib().orc_fft2(nx, ny, _p(buf), sign)
    return (buf[0::2] + 1j * buf[1::2]).reshape(ny, nx)


def ifft_c2r(spec: np.ndarray) -> np.ndarray:
    ny, nx = spec.shape
    re = np.ascontiguousarray(spec.real)
    im = np.ascontiguousarray(spec.imag)
    out = np.empty((ny, nx))
    lib().orc_ifft_c2r(nx, ny, _p(re), _p(im), _p(out))
    return out


def set_fp32_emulation_1d(on: bool):
    """Run the 1-D oracle's transforms in float arithmetic (a diagnostic of the
    FP32 device path's rounding; off by default)."""
    lib().orc1d_set_fp32_emulation(int(bool(on)))


def set_plant(code: int):
    """Switch on one planted mistake in the 2-D oracle (0 = none); see the
    list at the top of hawassi_oracle.c.  For tests that show a check fails
    on a plausible misreading."""
    lib().orc_set_plant(int(code))


def find_peaks(ec: np.ndarray, thr: float):
    """fun_peaks on the [ny, nx] field ec: list of (x, y) in the reference's order."""
    ec = np.ascontiguousarray(ec, dtype=np.float64)
    ny, nx = ec.shape
    xs = np.zeros(nx * ny, dtype=np.int32)
    ys = np.zeros(nx * ny, dtype=np.int32)
    n = lib().orc_find_peaks(_p(ec), nx, ny, float(thr), xs.ctypes.data, ys.ctypes.data)
    return list(zip(xs[:n].tolist(), ys[:n].tolist()))


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def num_threads() -> int:
    return lib().orc_num_threads()
