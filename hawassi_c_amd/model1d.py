"""Host-side mirror of the reference's 1-D driver around the HIP right-hand side.

``Hawassi1D`` owns one device batch (hw1d_create) of independent
realizations and exposes the reference's calling surface per realization:

* ``rhs(t, y, r=0)``       -- the GSL odeiv system callback rhs_HsF2
  (1d/rhs.c:96-241), 4-plane state [Re eta^][Im eta^][Re u^][Im u^];
* ``integrate(t1)``        -- ``while (t < ti) gsl_odeiv_evolve_apply`` of
  ode_solver_Ham (1d/functions.c:546-552) for every realization at once;

Errors raise ``HwError``; a missing library raises ``HipPathMissing`` --
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from .setup1d import PROPAGATION, Model1D


class Hawassi1D:
    def __init__(self, models: list[Model1D] | Model1D, precision: int = _lib.HW_PREC_FP32,
                 t0: float = 0.0, h0: float = 1e-4):
        if isinstance(models, Model1D):
            models = [models]
        self.lib = _lib.load()
        self.models = models
        self.precision = precision
        keep = []

        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return _lib.ptr(a)

        cfgs = (_lib.HwConfig1D * len(models))()
        for c, m in zip(cfgs, models):
            c.n, c.flat, c.source_on = m.n, int(m.flat), int(m.source_on)
            c.order = int(getattr(m, "order", 2))
            c.propagation = PROPAGATION[m.propagation]
            c.dx = m.dx
            c.k, c.aal, c.chi, c.damp, c.depth = arr(m.k), arr(m.aal), arr(m.chi), arr(m.damp), arr(m.depth)
            if m.flat:
                c.csqr, c.lop = arr(m.csqr), arr(m.lop)
            else:
                o = m.ops
                c.csq_min, c.csq_plus = arr(o["csq_min"]), arr(o["csq_plus"])
                c.gam_min, c.gam_plus = arr(o["gam_min"]), arr(o["gam_plus"])
                if o.get("gam_mid") is not None:
                    c.csq_mid, c.gam_mid = arr(o["csq_mid"]), arr(o["gam_mid"])
            if m.source_on:
                c.gamx_re, c.gamx_im = arr(m.gamx.real), arr(m.gamx.imag)
                c.gamx_skew_re, c.gamx_skew_im = arr(m.gamx_skew.real), arr(m.gamx_skew.imag)
                c.nt = len(m.t_sig)
                c.t_sig, c.sig, c.sig_skew = arr(m.t_sig), arr(m.sig), arr(m.sig_skew)
            br = m.breaking
            if br is not None:
                c.breaking_on = 1
                c.KBC, c.TC, c.Tstar, c.max_eta_init, c.delb = br["KBC"], br["TC"], br["Tstar"], br["MaxEtaInit"], br["delb"]
                for f in ("xbstart", "xbend", "xbendL", "xbstartR", "lenIF", "ILL", "IFF"):
                    setattr(c, f, int(br[f]))
            w = getattr(m, "wall", None)
            if w is not None:
                c.wall_on, c.idx_wall, c.refl_coef, c.t_wall0 = 1, int(w["idx"]), float(w["refl"]), float(w["t0"])
                c.gamw_skew_re, c.gamw_skew_im = arr(w["gamw_skew"].real), arr(w["gamw_skew"].imag)
            c.eps_abs, c.eps_rel = m.eps_abs, m.eps_rel
            c.t0, c.h0 = t0, h0
        h = ct.c_void_p()
        _lib.check(self.lib.hw1d_create(cfgs, len(models), precision, ct.byref(h)), "hw1d_create")
        self.ctx = h
        del keep

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.hw1d_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rhs(self, t: float, y: np.ndarray, r: int = 0) -> np.ndarray:
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.empty_like(y)
        _lib.check(self.lib.hw1d_rhs_at(self.ctx, r, float(t), _lib.ptr(y), _lib.ptr(out)), "hw1d_rhs_at")
        return out

    def set_state(self, y: np.ndarray, r: int = 0):
        y = np.ascontiguousarray(y, dtype=np.float64)
        _lib.check(self.lib.hw1d_set_state(self.ctx, r, _lib.ptr(y)), "hw1d_set_state")

    def get_state(self, r: int = 0) -> np.ndarray:
        y = np.empty(4 * self.models[r].n)
        _lib.check(self.lib.hw1d_get_state(self.ctx, r, _lib.ptr(y)), "hw1d_get_state")
        return y

    def set_time(self, t: float, h: float, r: int = 0):
        _lib.check(self.lib.hw1d_set_time(self.ctx, r, float(t), float(h)), "hw1d_set_time")

    def time(self, r: int = 0):
        t, h = ct.c_double(), ct.c_double()
        _lib.check(self.lib.hw1d_get_time(self.ctx, r, ct.byref(t), ct.byref(h)), "hw1d_get_time")
        return t.value, h.value

    def set_round_limit(self, max_rounds: int):
        """Bound the rounds one integrate call may take (0: none); past it the
        call raises HwError (HW_ERR_STEP) with the realizations where they got."""
        _lib.check(self.lib.hw1d_set_round_limit(self.ctx, int(max_rounds)), "hw1d_set_round_limit")

    def integrate(self, t1: float):
        _lib.check(self.lib.hw1d_integrate(self.ctx, float(t1)), "hw1d_integrate")

    def breaking_state(self, r: int = 0):
        B = np.empty(self.models[r].n)
        f = ct.c_int()
        _lib.check(self.lib.hw1d_get_breaking(self.ctx, r, _lib.ptr(B), ct.byref(f)), "hw1d_get_breaking")
        return B, f.value

    def stats(self, r: int = -1) -> dict:
        s = _lib.HwStats()
        _lib.check(self.lib.hw1d_get_stats(self.ctx, r, ct.byref(s)), "hw1d_get_stats")
        return dict(rhs_evals=s.rhs_evals, steps_accepted=s.steps_accepted, steps_rejected=s.steps_rejected,
                    rhs_ms_total=s.rhs_ms_total, rhs_timed=s.rhs_timed)
