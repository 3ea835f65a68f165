"""Host-side mirror of the reference's 2-D driver around the HIP right-hand side.

``Hawassi2D`` owns one device context (hw_ctx_create_2d) and exposes the
reference's calling surface:

* ``rhs(t, y, dydt)``            -- the GSL system callback (rhs_AB2_2DB_uv /
  rhs_AB2_2DF_uv, 2d/fun_odesolver.c:4458-4624), full 6-plane state;
* ``evolve_apply(t, t1, h)``     -- gsl_odeiv2_evolve_apply with rkf45 and
  control_y (2d/fun_odesolver.c:2567-2570, 2642);
* ``integrate(t, t1, h)``        -- the ``while (t < ti)`` loop of ode_solver_uv
  (2d/fun_odesolver.c:2641-2647);
* ``fields()``                   -- eta, u, v in physical space at the current
  state (2d/fun_odesolver.c:2663-2666).

Errors from the library raise ``HwError``; a missing library raises
``HipPathMissing`` -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from .setup2d import GRAV, Model2D


def initial_state(eta0: np.ndarray, u0: np.ndarray | None = None, v0: np.ndarray | None = None) -> np.ndarray:
    """set_init_ode (2d/fun_odesolver.c:1129-1161): forward transforms of the
    initial physical fields into the 6-plane full-grid state."""
    ny, nx = eta0.shape
    y = np.zeros(6 * nx * ny)
    planes = []
    for f in (eta0, u0, v0):
        s = np.fft.fft2(f) if f is not None else np.zeros((ny, nx), complex)
        planes += [s.real.ravel(), s.imag.ravel()]
    return np.concatenate(planes)


class Hawassi2D:
    """One device context of the 2-D path; with members > 1 it holds that many
    states of the model, advanced together (hw_ctx_create_2d_members):
    set_state / get_state then take / return arrays [members][6 nx ny] and
    integrate_members advances them.  full_grid=True (or a grid without
    fused kernels, or walls on a flat bottom) selects the full-grid path: the
    reference's own full complex state and transforms, any side 2^a 3^b 5^c."""

    def __init__(self, model: Model2D, precision: int = _lib.HW_PREC_FP32, members: int = 1, full_grid: bool = False):
        self.lib = _lib.load()
        self.members = members
        self.model = model
        self.precision = precision
        self.nx, self.ny = model.nx, model.ny
        self.dim = 6 * self.nx * self.ny
        keep = {}

        def arr(name, a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep[name] = a
            return _lib.ptr(a)

        c = _lib.HwConfig2D()
        c.nx, c.ny, c.precision, c.flat = self.nx, self.ny, precision, int(model.flat)
        c.grav, c.cpeak = GRAV, model.cpeak
        c.kx, c.ky = arr("kx", model.kx), arr("ky", model.ky)
        c.aal, c.fbdy = arr("aal", model.aal), arr("fbdy", model.fbdy)
        c.wallchar, c.chi = arr("wallchar", model.wallchar), arr("chi", model.chi)
        if model.flat:
            c.c2d, c.l2d = arr("c2d", model.c2d), arr("l2d", model.l2d)
            if model.wall:
                # fun_delphiH1hat's wall branches: the full-grid path
                c.wall, c.refl_one = 1, int(model.refl_one)
                if model.refl_one:
                    c.grad_wall_x, c.grad_wall_y = arr("gwx", model.grad_wall_x), arr("gwy", model.grad_wall_y)
        else:
            o = model.ops
            c.gam_min, c.gam_plus = arr("gm", o["gam_min"]), arr("gp", o["gam_plus"])
            c.csq_min, c.csq_plus = arr("cm", o["csq_min"]), arr("cp", o["csq_plus"])
            c.om2_min, c.om2_plus = arr("om", o["om2_min"]), arr("op", o["om2_plus"])
            if o.get("gam_mid") is not None:
                c.gam_mid, c.csq_mid, c.om2_mid = arr("gc", o["gam_mid"]), arr("cc", o["csq_mid"]), arr("oc", o["om2_mid"])
        c.source_on = int(model.source_on)
        if model.source_on:
            c.vertical, c.propagation = int(model.vertical), int(model.propagation)
            c.new_gam, c.skew_gam = arr("ng", model.new_gam), arr("sg", model.skew_gam)
            c.nline, c.nt = model.sig.shape
            c.t_sig, c.sig, c.skew_sig = arr("ts", model.t_sig), arr("sig", model.sig), arr("sk", model.skew_sig)
        c.eps_abs, c.eps_rel = model.eps_abs, model.eps_rel
        br = model.breaking
        if br is not None:
            c.breaking_on = 1
            c.KBC, c.TC, c.Tchar, c.delb = br["KBC"], br["TC"], br["Tchar"], br["delb"]
            c.max_eta_init, c.tstar, c.rsearch, c.current = br["MaxEtaInit"], br["Tstar"], br["Rsearch"], 0.0
            c.pbchar = arr("pbchar", br["char"])
            c.depth = arr("depth", model.depth)
            c.x, c.y = arr("x", model.x), arr("y", model.y)
        c.t0 = float(model.meta.get("t0", 0.0))
        c.model = int(model.order)
        # the full-grid path: asked for, a grid without fused kernels, or walls on a flat bottom
        self.full_grid = bool(full_grid or (model.flat and model.wall) or
                              not self.lib.hw_supported(self.nx, self.ny, precision))
        c.full_grid = int(self.full_grid)
        self._cfg, self._keep = c, keep
        h = ct.c_void_p()
        if members == 1:
            _lib.check(self.lib.hw_ctx_create_2d(ct.byref(c), ct.byref(h)), "hw_ctx_create_2d")
        else:
            _lib.check(self.lib.hw_ctx_create_2d_members(ct.byref(c), int(members), ct.byref(h)),
                       "hw_ctx_create_2d_members")
        self.ctx = h
        self._keep = None   # data now lives on the device

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.hw_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------- GSL-style surface --
    def rhs(self, t: float, y: np.ndarray, dydt: np.ndarray | None = None) -> np.ndarray:
        y = np.ascontiguousarray(y, dtype=np.float64)
        if dydt is None:
            dydt = np.empty_like(y)
        _lib.check(self.lib.hw_rhs(float(t), _lib.ptr(y), _lib.ptr(dydt), self.ctx), "hw_rhs")
        return dydt

    def set_state(self, y: np.ndarray):
        y = np.ascontiguousarray(y, dtype=np.float64)
        if y.size != self.members * self.dim:
            raise ValueError(f"state has {y.size} values, the context holds {self.members} x {self.dim}")
        _lib.check(self.lib.hw_set_state(self.ctx, _lib.ptr(y)), "hw_set_state")

    def get_state(self) -> np.ndarray:
        y = np.empty(self.members * self.dim)
        _lib.check(self.lib.hw_get_state(self.ctx, _lib.ptr(y)), "hw_get_state")
        return y if self.members == 1 else y.reshape(self.members, self.dim)

    def integrate_members(self, t, t1: float, h):
        """Advance every member to t1 (hw_integrate_members); t and h per
        member; returns the new (t, h) lists."""
        tt = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (self.members,)))
        hh = np.ascontiguousarray(np.broadcast_to(np.asarray(h, dtype=np.float64), (self.members,)))
        _lib.check(self.lib.hw_integrate_members(self.ctx, _lib.ptr(tt), float(t1), _lib.ptr(hh)),
                   "hw_integrate_members")
        return list(tt), list(hh)

    def evolve_apply(self, t: float, t1: float, h: float):
        tt, hh = ct.c_double(t), ct.c_double(h)
        _lib.check(self.lib.hw_evolve_apply(self.ctx, ct.byref(tt), float(t1), ct.byref(hh)), "hw_evolve_apply")
        return tt.value, hh.value

    def integrate(self, t: float, t1: float, h: float):
        tt, hh = ct.c_double(t), ct.c_double(h)
        _lib.check(self.lib.hw_integrate(self.ctx, ct.byref(tt), float(t1), ct.byref(hh)), "hw_integrate")
        return tt.value, hh.value

    def fields(self):
        n = self.nx * self.ny
        eta, u, v = np.empty(n), np.empty(n), np.empty(n)
        _lib.check(self.lib.hw_get_fields(self.ctx, _lib.ptr(eta), _lib.ptr(u), _lib.ptr(v)), "hw_get_fields")
        sh = (self.ny, self.nx)
        return eta.reshape(sh), u.reshape(sh), v.reshape(sh)

    def stats(self) -> dict:
        s = _lib.HwStats()
        _lib.check(self.lib.hw_get_stats(self.ctx, ct.byref(s)), "hw_get_stats")
        return dict(rhs_evals=s.rhs_evals, steps_accepted=s.steps_accepted, steps_rejected=s.steps_rejected,
                    rhs_ms_total=s.rhs_ms_total, rhs_timed=s.rhs_timed)

    def set_tprev(self, t: float):
        _lib.check(self.lib.hw_set_tprev(self.ctx, float(t)), "hw_set_tprev")

    def breaking_state(self):
        B = np.empty(self.nx * self.ny)
        f = ct.c_int()
        _lib.check(self.lib.hw_get_breaking(self.ctx, _lib.ptr(B), ct.byref(f)), "hw_get_breaking")
        return B.reshape(self.ny, self.nx), f.value

    def reset_stats(self):
        _lib.check(self.lib.hw_reset_stats(self.ctx), "hw_reset_stats")

    def set_step_mode(self, graph: bool):
        """integrate / integrate_batch with the step control on the device and
        each attempted step replayed from a HIP graph (True), or on the host
        (False, the default); see hw_set_step_mode."""
        _lib.check(self.lib.hw_set_step_mode(self.ctx, _lib.HW_STEP_GRAPH if graph else _lib.HW_STEP_HOST),
                   "hw_set_step_mode")

    def sync(self):
        _lib.check(self.lib.hw_sync(self.ctx), "hw_sync")


def integrate_batch(solvers: "list[Hawassi2D]", t: "list[float]", t1: float, h: "list[float]"):
    """Advance several contexts to t1 in lockstep (hw_integrate_batch).

    Each context keeps its own step control; the result equals calling
    integrate on each alone.  Returns the new (t, h) lists.  When every
    context steps on the device (set_step_mode(True)), contexts with several
    members take part, and t and h hold one value per member.
    """
    n = len(solvers)
    if n == 0:
        return [], []
    if len(t) != sum(s.members for s in solvers) or len(h) != len(t):
        raise ValueError("t and h need one value per member of every context")
    lib = solvers[0].lib
    arr = (ct.c_void_p * n)(*[s.ctx for s in solvers])
    tt = np.ascontiguousarray(t, dtype=np.float64).copy()
    hh = np.ascontiguousarray(h, dtype=np.float64).copy()
    _lib.check(lib.hw_integrate_batch(arr, n, _lib.ptr(tt), float(t1), _lib.ptr(hh)), "hw_integrate_batch")
    return tt.tolist(), hh.tolist()


def fft2_forward(phys: np.ndarray, precision: int = _lib.HW_PREC_FP64):
    """Device forward transform of a real [ny, nx] field (full-grid result)."""
    lib = _lib.load()
    ny, nx = phys.shape
    p = np.ascontiguousarray(phys, dtype=np.float64)
    re, im = np.empty_like(p), np.empty_like(p)
    _lib.check(lib.hw_fft2_forward(nx, ny, precision, _lib.ptr(p), _lib.ptr(re), _lib.ptr(im)), "hw_fft2_forward")
    return re + 1j * im


def fft2_inverse(spec: np.ndarray, precision: int = _lib.HW_PREC_FP64):
    """Device inverse with the reference's projection semantics; also returns
    Im of the Nyquist column per row (scaled by 1/N)."""
    lib = _lib.load()
    ny, nx = spec.shape
    re = np.ascontiguousarray(spec.real, dtype=np.float64)
    im = np.ascontiguousarray(spec.imag, dtype=np.float64)
    out = np.empty((ny, nx))
    nyq = np.empty(ny)
    _lib.check(lib.hw_fft2_inverse(nx, ny, precision, _lib.ptr(re), _lib.ptr(im), _lib.ptr(out), _lib.ptr(nyq)),
               "hw_fft2_inverse")
    return out, nyq
