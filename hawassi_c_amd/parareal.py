"""Learned parareal on the 2-D path: a spectral neural operator (fno.FNO) as
the coarse propagator G, the plain device solver (model2d.Hawassi2D, FP32
transforms with FP64 state and stage sums) as the fine propagator F,
corrector and judge, and the serial fine solver as the fallback.

The reference has no time parallelism: ode_solver_uv integrates serially
(2d/fun_odesolver.c:2565-2872).  This module adds it for autonomous problems
(no influx, breaking off: workloads.shoal), where every slice can be advanced
over [0, dt] in its own device context and u = grad phi holds, so the
parareal variable is the pair of physical fields (eta, phi).

Parareal (Lions, Maday and Turinici 2001) with slices n = 0..N-1:
    U^0_{n+1} = G(U^0_n);
    U^{k+1}_{n+1} = G(U^{k+1}_n) + F(U^k_n) - G(U^k_n),
converged when the largest relative change of a slice between iterations is
at most `tol`.  After k iterations the first k slices equal the serial fine
solution exactly, so iteration k advances only slices n >= k with F (all of
them together, hw_integrate_batch).  Fallback: when the change stalls (does
not halve over two iterations) or exceeds 1, when a slice's field norm grows
beyond grow_limit times the initial one, or when the iteration budget ends,
the remaining slices are advanced serially with F from the first one not yet
converged.
"""
from __future__ import annotations

import math
import time

import numpy as np

import scipy.fft as _fft

_W = {"workers": 16}       # host threads of the field conversions

from . import _lib
from .model2d import Hawassi2D, integrate_batch


class Codec:
    """(eta, phi) physical <-> the 6-plane full-grid state of the 2-D path
    (u = d phi/dx, v = d phi/dy with the model's wavenumbers)."""

    def __init__(self, model):
        self.m = model
        n = model.nx
        self.KX = model.kx[None, :]
        self.KY = model.ky[:, None]
        k2 = self.KX ** 2 + self.KY ** 2
        self.inv_k2 = np.where(k2 > 1e-12, 1.0 / np.where(k2 > 1e-12, k2, 1.0), 0.0)
        self.n = n

    def to_state(self, eta, phi):
        """u^ = i kx phi^, v^ = i ky phi^ on the stored columns kx = 0..nx/2
        (the only ones the device reads), the redundant half as their mirror."""
        n = self.n
        h = n // 2 + 1
        eh = _fft.rfft2(eta, **_W)
        ph = _fft.rfft2(phi, **_W)
        uh = 1j * self.KX[:, :h] * ph
        vh = 1j * self.KY * ph
        planes = []
        for s in (eh, uh, vh):
            full = np.empty((n, n), dtype=complex)
            full[:, :h] = s
            full[:, h:] = np.conj(s[(-np.arange(n)) % n][:, 1:n - h + 1][:, ::-1])
            planes += [full.real.ravel(), full.imag.ravel()]
        return np.concatenate(planes)

    def from_state(self, y):
        n = self.n
        N = n * n
        pl = [(y[2 * f * N:(2 * f + 1) * N] + 1j * y[(2 * f + 1) * N:(2 * f + 2) * N]).reshape(n, n)
              for f in range(3)]
        h = n // 2 + 1
        eta = _half_inverse(pl[0][:, :h], n)
        ph = -1j * (self.KX[:, :h] * pl[1][:, :h] + self.KY * pl[2][:, :h]) * self.inv_k2[:, :h]
        phi = _half_inverse(ph, n)
        return eta, phi


def _half_inverse(s, n):
    """The reference's inverse (Hermitian projection of column 0, real part)
    of the stored half spectrum [ny, nx/2 + 1]."""
    s = s.copy()
    s[n // 2 + 1:, 0] = np.conj(s[1: n // 2][::-1, 0])
    s[0, 0] = s[0, 0].real
    return _fft.irfft2(s, s=(n, n), **_W)


class LinearFlat:
    """The coarse propagator without the network: exact linear propagation
    over dt on a flat bottom of depth d0 (the basin's far-field depth), mode by
    mode: eta^(dt) = eta^ cos(w dt) + (w/g) phi^ sin(w dt),
    phi^(dt) = phi^ cos(w dt) - (g/w) eta^ sin(w dt), w = sqrt(g k tanh k d0),
    inside the model's anti-alias band; outside it the fine solver's
    right-hand side vanishes (it is multiplied by the band mask), so those
    modes are carried over unchanged."""

    def __init__(self, model, dt: float, d0: float | None = None):
        from .setup2d import GRAV
        n = model.nx
        d0 = float(model.depth.max()) if d0 is None else d0
        # the model's own wavenumbers on the stored columns kx = 0..nx/2
        K = np.hypot(np.asarray(model.kx)[None, :n // 2 + 1], np.asarray(model.ky)[:, None])
        w = np.sqrt(GRAV * K * np.tanh(K * d0))
        self.c, self.s = np.cos(w * dt), np.sin(w * dt)
        self.a = np.where(K > 0, w / GRAV, 0.0)
        self.b = np.where(K > 0, GRAV / np.where(w > 0, w, 1.0), 0.0)
        self.mask = np.asarray(model.aal, dtype=float)[:, :n // 2 + 1]
        self.n = n

    def spectral(self, eh, ph):
        e2 = eh * self.c + self.a * ph * self.s
        p2 = ph * self.c - self.b * eh * self.s
        p2[0, 0] = 0.0
        return e2, p2

    def __call__(self, eta, phi):
        eh, ph = _fft.rfft2(eta, **_W), _fft.rfft2(phi, **_W)
        e2, p2 = self.spectral(eh, ph)
        keep = 1.0 - self.mask
        return (_fft.irfft2(self.mask * e2 + keep * eh, s=(self.n, self.n), **_W),
                _fft.irfft2(self.mask * p2 + keep * ph, s=(self.n, self.n), **_W))


class LinearBlend(LinearFlat):
    """Linear propagation over dt through a slowly varying depth: the flat
    propagators of `ndepth` reference depths spanning the basin's depth
    range, each applied to the whole field, blended at every point linearly
    in the local depth (the local-phase-speed approximation of refraction
    over a shoal; exact on a flat bottom)."""

    def __init__(self, model, dt: float, ndepth: int = 4):
        dmin, dmax = float(model.depth.min()), float(model.depth.max())
        self.depths = np.linspace(dmin, dmax, ndepth) if dmax > dmin else np.array([dmax])
        self.parts = [LinearFlat(model, dt, d) for d in self.depths]
        super().__init__(model, dt, dmax)
        D = np.asarray(model.depth, dtype=float)
        self.wts = []
        nd = len(self.depths)
        if nd == 1:
            self.wts = [np.ones_like(D)]
        else:
            pos = np.clip((D - dmin) / (dmax - dmin) * (nd - 1), 0.0, nd - 1)
            for j in range(nd):
                self.wts.append(np.clip(1.0 - np.abs(pos - j), 0.0, 1.0))

    def __call__(self, eta, phi):
        n = self.n
        eh, ph = _fft.rfft2(eta, **_W), _fft.rfft2(phi, **_W)
        keep = 1.0 - self.mask
        eta2 = _fft.irfft2(keep * eh, s=(n, n), **_W)
        phi2 = _fft.irfft2(keep * ph, s=(n, n), **_W)
        for part, w in zip(self.parts, self.wts):
            e2, p2 = part.spectral(eh, ph)
            eta2 += w * _fft.irfft2(self.mask * e2, s=(n, n), **_W)
            phi2 += w * _fft.irfft2(self.mask * p2, s=(n, n), **_W)
        return eta2, phi2


class CoarseFNO:
    """G(eta, phi) = base(eta, phi) + scale * FNO(normalised [eta, phi, depth]):
    the network learns what the base propagator (LinearFlat, LinearBlend)
    misses over one coarse step.  Its part is kept to the anti-alias band and
    to |k| below twice its retained modes (the band of its spectral layers
    and their quadratic harmonics): what its pointwise layers put above that
    is noise, and noise at high wavenumbers makes the fine solver's steps
    small and, iterated over many slices, grows."""

    def __init__(self, net, depth, stats, base):
        # parareal differences G(U^{k+1}) - G(U^k) over states that converge:
        # BF16 products would leave a rounding floor of ~1e-3 in them
        net.set_precision(True)
        self.net = net
        self.depth = depth
        self.st = stats
        self.base = base
        n = base.n
        self.band = base.mask.copy()
        modes = getattr(net, "modes", None)
        if modes is not None:
            ix = np.fft.rfftfreq(n) * n
            iy = np.fft.fftfreq(n) * n
            self.band = self.band * (np.hypot(ix[None, :], iy[:, None]) < 2 * modes)

    def inputs(self, eta, phi):
        s = self.st
        return np.stack([(eta - s["mu"][0]) / s["sd"][0], (phi - s["mu"][1]) / s["sd"][1],
                         (self.depth - s["mu"][2]) / s["sd"][2]]).astype(np.float32)

    def __call__(self, eta, phi):
        out = self.net.forward(self.inputs(eta, phi)[None])[0].astype(np.float64)
        b = self.base
        e0, p0 = b(eta, phi)
        ce = _fft.irfft2(self.band * _fft.rfft2(self.st["dsd"][0] * out[0], **_W), s=(b.n, b.n), **_W)
        cp = _fft.irfft2(self.band * _fft.rfft2(self.st["dsd"][1] * out[1], **_W), s=(b.n, b.n), **_W)
        return e0 + ce, p0 + cp


def training_seas(model, seeds, hs=(0.01, 0.06), tp: float = 2.0):
    """Initial seas of the training trajectories: per seed a significant
    height drawn log-uniformly from `hs` and a mean direction drawn
    uniformly, so that the data covers every direction of travel over the
    shoal (workloads.shoal_sea)."""
    from . import workloads as W
    out = []
    for s in seeds:
        g = np.random.default_rng([0x5EA, s])
        h = float(np.exp(g.uniform(math.log(hs[0]), math.log(hs[1]))))
        out.append(W.shoal_sea(model, s, hs=h, tp=tp, theta0=float(g.uniform(0, 2 * math.pi))))
    return out


def snapshots(model, initial, t_max: float, dt_snap: float = 1.0, precision: int = _lib.HW_PREC_FP32,
              h0: float = 0.05, log=None):
    """Trajectories of the fine solver from the initial (eta, phi) fields
    `initial` (e.g. training_seas), all advanced together; returns the
    fields [trajectory][time][2][ny][nx] (float32) at t = 0, dt_snap, ...,
    t_max and the times."""
    codec = Codec(model)
    ctx = [Hawassi2D(model, precision) for _ in initial]
    out = []
    for c, (e, p) in zip(ctx, initial):
        c.set_state(codec.to_state(e, p))
        out.append([np.stack([e, p]).astype(np.float32)])
    times = [0.0]
    t = [0.0] * len(ctx)
    h = [h0] * len(ctx)
    w0 = time.perf_counter()
    nsnap = int(round(t_max / dt_snap))
    for i in range(1, nsnap + 1):
        t1 = i * dt_snap
        t, h = integrate_batch(ctx, t, t1, h)
        for j, c in enumerate(ctx):
            e, p = codec.from_state(c.get_state())
            out[j].append(np.stack([e, p]).astype(np.float32))
        times.append(t1)
        if log:
            log(f"snapshots t={t1:.1f} wall {time.perf_counter() - w0:.1f} s")
    for c in ctx:
        c.close()
    return np.asarray(out), np.asarray(times)


def training_pairs(traj, depth, dt_coarse: float, base, dt_snap: float = 1.0):
    """(input, target) pairs (t, t + dt_coarse) from trajectories
    [seed][time][2][ny][nx]: the normalised arrays X [pairs][3][ny][nx],
    Y [pairs][2][ny][nx] (float32) and the normalisation statistics; the
    target is what `base` misses of the fine solver's step."""
    lag = int(round(dt_coarse / dt_snap))
    a = traj[:, :-lag].reshape(-1, 2, *traj.shape[-2:])
    b = traj[:, lag:].reshape(-1, 2, *traj.shape[-2:])
    d = np.empty_like(b)
    for i in range(a.shape[0]):
        e0, p0 = base(a[i, 0].astype(np.float64), a[i, 1].astype(np.float64))
        d[i, 0] = b[i, 0] - e0
        d[i, 1] = b[i, 1] - p0
    mu = [float(a[:, 0].mean()), float(a[:, 1].mean()), float(depth.mean())]
    sd = [float(a[:, 0].std()), float(a[:, 1].std()), float(max(depth.std(), 1e-12))]
    dsd = [float(d[:, 0].std()), float(d[:, 1].std())]
    st = {"mu": mu, "sd": sd, "dsd": dsd}
    X = np.empty((a.shape[0], 3) + a.shape[-2:], dtype=np.float32)
    X[:, 0] = (a[:, 0] - mu[0]) / sd[0]
    X[:, 1] = (a[:, 1] - mu[1]) / sd[1]
    X[:, 2] = ((depth - mu[2]) / sd[2]).astype(np.float32)
    Y = np.empty((a.shape[0], 2) + a.shape[-2:], dtype=np.float32)
    Y[:, 0] = d[:, 0] / dsd[0]
    Y[:, 1] = d[:, 1] / dsd[1]
    return X, Y, st


def train(net, X, Y, epochs: int, batch: int, lr: float, seed: int = 0, log=None, holdout: int = 0):
    """Adam on the device (cosine learning-rate decay); returns the mean
    training loss of each epoch and, with a hold-out, its final loss."""
    net.set_dataset(X, Y)
    g = np.random.default_rng(seed)
    ntr = X.shape[0] - holdout
    hist = []
    steps = max(1, ntr // batch)
    total = epochs * steps
    it = 0
    w0 = time.perf_counter()
    for ep in range(epochs):
        perm = g.permutation(ntr)
        acc = 0.0
        for s in range(steps):
            lr_t = 0.5 * lr * (1 + math.cos(math.pi * it / total))
            acc += net.train_step(perm[s * batch:(s + 1) * batch], lr_t)
            it += 1
        hist.append(acc / steps)
        if log:
            log(f"epoch {ep + 1}/{epochs} loss {hist[-1]:.4e} wall {time.perf_counter() - w0:.1f} s")
    val = None
    if holdout:
        val = float(np.mean([net.loss_grad(X[i:i + 1], Y[i:i + 1])[0] for i in range(ntr, X.shape[0])]))
    return hist, val


class Parareal:
    """Parareal over `nslices` slices of length `dt` for the autonomous
    problem `model`.  The fine propagator of an iteration advances every
    slice together: by default one context holding the slices as members,
    every kernel covering all of them, step control on the device and each
    attempted step a HIP-graph replay (hw_ctx_create_2d_members); the serial
    fine solver is a single context stepping the same way, so batched and
    serial slices agree bit for bit.  device_steps=False uses one context per
    slice in lockstep with the step control on the host (hw_integrate_batch)."""

    def __init__(self, model, coarse, nslices: int, dt: float, precision: int = _lib.HW_PREC_FP32, h0: float = 0.05,
                 grow_limit: float = 4.0, device_steps: bool = True):
        self.m, self.G, self.N, self.dt, self.h0 = model, coarse, nslices, dt, h0
        # the basin conserves energy: a slice whose field norm exceeds
        # grow_limit times the initial one has diverged, and the fine solver
        # would crawl through its noise
        self.grow_limit = grow_limit
        self.codec = Codec(model)
        self.precision = precision
        self.device_steps = device_steps
        self.single = Hawassi2D(model, precision)
        if device_steps:
            self.single.set_step_mode(True)
            self.ctx = []
            self.members = {}             # slice count -> context holding that many members
        else:
            self.ctx = [self.single] + [Hawassi2D(model, precision) for _ in range(nslices - 1)]
        self.fine_rhs = 0

    def close(self):
        for c in self.ctx:
            c.close()
        if self.device_steps:
            for c in self.members.values():
                c.close()
            self.members = {}
            self.single.close()
        self.ctx = []

    def _members(self, count):
        if count not in self.members:
            self.members[count] = Hawassi2D(self.m, self.precision, members=count)
        return self.members[count]

    def fine(self, states, first: int = 0):
        """F(U_n) for n >= first, all together; returns the list (None below first)."""
        out = [None] * self.N
        if self.device_steps:
            c = self._members(self.N)
            ys = np.zeros((self.N, c.dim))
            for i in range(first, self.N):
                ys[i] = self.codec.to_state(*states[i])
            c.set_state(ys)
            c.reset_stats()
            # slices below `first` start at t1: they do not move
            c.integrate_members([self.dt if i < first else 0.0 for i in range(self.N)], self.dt, self.h0)
            Y = c.get_state()
            for i in range(first, self.N):
                out[i] = self.codec.from_state(Y[i])
            self.fine_rhs += c.stats()["rhs_evals"]
            return out
        act = list(range(first, self.N))
        cs = [self.ctx[i] for i in act]
        for i, c in zip(act, cs):
            c.set_state(self.codec.to_state(*states[i]))
            c.reset_stats()
        integrate_batch(cs, [0.0] * len(cs), self.dt, [self.h0] * len(cs))
        for i, c in zip(act, cs):
            out[i] = self.codec.from_state(c.get_state())
            self.fine_rhs += c.stats()["rhs_evals"]
        return out

    def fine_serial(self, u0, first: int = 0, start=None):
        """The plain fine solver, slice after slice (one context)."""
        c = self.single
        U = [None] * (self.N + 1)
        U[first] = u0 if start is None else start
        for n in range(first, self.N):
            c.set_state(self.codec.to_state(*U[n]))
            c.integrate(0.0, self.dt, self.h0)
            U[n + 1] = self.codec.from_state(c.get_state())
        return U

    def grown(self, U, u0, first):
        ref = [max(float(np.linalg.norm(u0[f])), 1e-300) for f in range(2)]
        return any(float(np.linalg.norm(U[n][f])) > self.grow_limit * ref[f]
                   for n in range(first, len(U)) for f in range(2))

    @staticmethod
    def change(a, b):
        return max(float(np.linalg.norm(a[f] - b[f]) / max(np.linalg.norm(a[f]), 1e-300)) for f in range(2))

    def run_windows(self, u0, window: int, tol: float = 1e-4, log=None, reference=None):
        """Parareal on consecutive windows of `window` slices (each window
        starts from the converged end of the previous one): the coarse error
        grows along a sweep (after one iteration it is about the window length
        times the one-step error), so a wave problem needs windows short
        enough that the iteration contracts.  Returns the run() fields summed
        over the windows."""
        N, W = self.N, window
        U = [u0] + [None] * N
        tot = {"iterations": 0, "iterations_per_window": [], "change": [], "error": [], "fallback_windows": 0,
               "wall_s": 0.0, "coarse_s": 0.0, "fine_s": 0.0}
        saveN = self.N
        for w0 in range(0, N, W):
            nw = min(W, N - w0)
            self.N = nw
            ref = None if reference is None else reference[w0:w0 + nw + 1]
            r = self.run(U[w0], tol=tol, max_iter=nw, reference=ref, log=log)
            for i in range(1, nw + 1):
                U[w0 + i] = r["U"][i]
            tot["iterations"] += r["iterations"]
            tot["iterations_per_window"].append(r["iterations"])
            tot["change"].append(r["change"])
            tot["error"].append(r["error"])
            tot["fallback_windows"] += int(r["fallback"])
            for k in ("wall_s", "coarse_s", "fine_s"):
                tot[k] += r[k]
            if log:
                log(f"window {w0 // W + 1}: {r['iterations']} iterations, change {r['change'][-1]:.2e}"
                    + (f" error {r['error'][-1]:.2e}" if r["error"] else "") + (" (fallback)" if r["fallback"] else ""))
        self.N = saveN
        tot["U"] = U
        if reference is not None:
            tot["final_error"] = max(self.change(reference[n], U[n]) for n in range(1, N + 1))
        return tot

    def run(self, u0, tol: float = 1e-4, max_iter: int | None = None, log=None, reference=None):
        """Returns a dict: slices U[0..N], iterations, the change and (given
        the serial fine solution `reference`) the error per iteration,
        whether the fallback ran, and the wall-clock split."""
        N = self.N
        max_iter = N if max_iter is None else max_iter
        t0 = time.perf_counter()
        tc = tf = 0.0
        U = [u0] + [None] * N
        Gold = [None] * N
        w = time.perf_counter()
        for n in range(N):
            Gold[n] = self.G(*U[n])
            U[n + 1] = Gold[n]
        tc += time.perf_counter() - w
        hist, errs = [], []
        fallback = False
        k = 0
        converged = False
        if self.grown(U, u0, 1):
            max_iter = 0
            if log:
                log("parareal: the coarse sweep diverged")
        while k < max_iter:
            w = time.perf_counter()
            Fk = self.fine(U, first=k)
            tf += time.perf_counter() - w
            w = time.perf_counter()
            Unew = U[:k + 1] + [None] * (N - k)
            for n in range(k, N):
                g_new = self.G(*Unew[n])
                Unew[n + 1] = tuple(g_new[f] + Fk[n][f] - Gold[n][f] for f in range(2))
                Gold[n] = g_new
            tc += time.perf_counter() - w
            ch = max(self.change(Unew[n], U[n]) for n in range(k + 1, N + 1))
            U = Unew
            k += 1
            hist.append(ch)
            if reference is not None:
                errs.append(max(self.change(reference[n], U[n]) for n in range(1, N + 1)))
            if log:
                log(f"parareal iteration {k}: change {ch:.3e}" + (f" error {errs[-1]:.3e}" if errs else ""))
            if ch <= tol:
                converged = True
                break
            # stall or divergence: hand over to the serial fine solver
            if not np.isfinite(ch) or ch > 1.0 or (k >= 3 and hist[-1] > 0.5 * hist[-3]) or self.grown(U, u0, k + 1):
                break
        if not converged and k >= N:
            converged = True        # every slice has been advanced by F from an exact slice
        if not converged:
            # fallback: the plain fine solver from the first slice that has not converged
            fallback = True
            w = time.perf_counter()
            Us = self.fine_serial(None, first=k, start=U[k])
            for n in range(k + 1, N + 1):
                U[n] = Us[n]
            tf += time.perf_counter() - w
        return {"U": U, "iterations": k, "change": hist, "error": errs, "fallback": fallback,
                "converged": converged, "wall_s": time.perf_counter() - t0, "coarse_s": tc, "fine_s": tf}
