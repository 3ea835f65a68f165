"""TEST INFRASTRUCTURE ONLY -- FP64 numpy restatement of the spectral neural
operator of hawassi_c_amd/csrc/fno.hip (forward pass, mean-squared-error loss
and the hand-written backward pass), the checker of the device's BF16-MFMA
kernels.  The operator has no counterpart in the reference.

Conventions (the device's): parameters in one flat vector (lift W [C][cin],
b [C]; per layer W [C][C], b [C], spectral weights R [2][C_in][C_out][2m][m];
projection W1 [H][C], b1 [H], W2 [cout][H], b2 [cout]); fields [batch][ch][n][n];
retained modes kx in [0, m), ky in [0, m) and [n-m, n) (row index i of R's
mode axis: ky = i for i < m, n - 2m + i otherwise); forward transform
exp(-i), inverse Re of the truncated backward transform / n^2 with kx > 0
counted twice; GELU 0.5 x (1 + erf(x / sqrt 2)) after every Fourier layer but
the last and on the projection's hidden layer.

The backward pass is pinned by central differences in tests/test_fno.py.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


class FNOOracle:
    def __init__(self, n, width, modes, layers, hidden, cin=3, cout=2):
        self.n, self.C, self.m, self.L, self.H, self.cin, self.cout = n, width, modes, layers, hidden, cin, cout
        self.kys = np.r_[np.arange(modes), np.arange(n - modes, n)]
        C, m = width, modes
        sizes = [("pw", (C, cin)), ("pb", (C,))]
        for l in range(layers):
            sizes += [(f"w{l}", (C, C)), (f"b{l}", (C,)), (f"r{l}", (2, C, C, 2 * m, m))]
        sizes += [("q1w", (hidden, C)), ("q1b", (hidden,)), ("q2w", (cout, hidden)), ("q2b", (cout,))]
        self.layout, off = [], 0
        for name, shp in sizes:
            cnt = int(np.prod(shp))
            self.layout.append((name, off, shp))
            off += cnt
        self.nparam = off

    def unpack(self, p):
        return {name: np.asarray(p[o:o + int(np.prod(shp))], dtype=np.float64).reshape(shp)
                for name, o, shp in self.layout}

    def pack(self, d):
        out = np.zeros(self.nparam)
        for name, o, shp in self.layout:
            out[o:o + int(np.prod(shp))] = d[name].ravel()
        return out

    # ------------------------------------------------------------- spectral
    def _spec_fwd(self, v, R):
        """v [b][C][n][n] -> (zh [b][C][2m][m] complex, y [b][C][n][n])."""
        m, n = self.m, self.n
        full = np.fft.fft2(v)                               # exp(-i), unnormalised
        zh = full[:, :, self.kys, :m]
        Rc = R[0] + 1j * R[1]                               # [i][o][2m][m]
        yh = np.einsum("bimk,iomk->bomk", zh, Rc)
        S = np.zeros(v.shape, dtype=complex)
        w = np.where(np.arange(m) == 0, 1.0, 2.0)
        S[:, :, self.kys, :m] = yh * w[None, None, None, :]
        y = np.real(np.fft.ifft2(S))
        return zh, yh, y

    def _spec_bwd(self, dy, zh, R):
        m, n = self.m, self.n
        w = np.where(np.arange(m) == 0, 1.0, 2.0)
        g_yh = np.fft.fft2(dy)[:, :, self.kys, :m] * w[None, None, None, :] / (n * n)
        Rc = R[0] + 1j * R[1]
        dR = np.einsum("bimk,bomk->iomk", np.conj(zh), g_yh)
        g_zh = np.einsum("bomk,iomk->bimk", g_yh, np.conj(Rc))
        G = np.zeros(dy.shape, dtype=complex)
        G[:, :, self.kys, :m] = g_zh
        dv = np.real(np.fft.ifft2(G)) * (n * n)
        return dv, np.stack([dR.real, dR.imag])

    # ---------------------------------------------------------------- model
    def forward(self, p, x):
        P = self.unpack(p)
        cache = {"x": x}
        v = np.einsum("ci,bipq->bcpq", P["pw"], x) + P["pb"][None, :, None, None]
        cache["v0"] = v
        for l in range(self.L):
            zh, yh, s = self._spec_fwd(v, P[f"r{l}"])
            pre = s + np.einsum("oi,bipq->bopq", P[f"w{l}"], v) + P[f"b{l}"][None, :, None, None]
            cache[f"zh{l}"], cache[f"pre{l}"], cache[f"v{l}"] = zh, pre, v
            v = gelu(pre) if l < self.L - 1 else pre
        cache["vL"] = v
        hpre = np.einsum("hc,bcpq->bhpq", P["q1w"], v) + P["q1b"][None, :, None, None]
        h = gelu(hpre)
        out = np.einsum("oh,bhpq->bopq", P["q2w"], h) + P["q2b"][None, :, None, None]
        cache["hpre"], cache["h"] = hpre, h
        return out, cache

    def loss_grad(self, p, x, t):
        P = self.unpack(p)
        out, c = self.forward(p, x)
        cnt = out.size
        loss = float(np.mean((out - t) ** 2))
        d = 2.0 * (out - t) / cnt
        g = {}
        g["q2b"] = d.sum(axis=(0, 2, 3))
        g["q2w"] = np.einsum("bopq,bhpq->oh", d, c["h"])
        dh = np.einsum("oh,bopq->bhpq", P["q2w"], d) * gelu_d(c["hpre"])
        g["q1b"] = dh.sum(axis=(0, 2, 3))
        g["q1w"] = np.einsum("bhpq,bcpq->hc", dh, c["vL"])
        dv = np.einsum("hc,bhpq->bcpq", P["q1w"], dh)
        for l in range(self.L - 1, -1, -1):
            dpre = dv * gelu_d(c[f"pre{l}"]) if l < self.L - 1 else dv
            g[f"b{l}"] = dpre.sum(axis=(0, 2, 3))
            g[f"w{l}"] = np.einsum("bopq,bipq->oi", dpre, c[f"v{l}"])
            dv_w = np.einsum("oi,bopq->bipq", P[f"w{l}"], dpre)
            dv_s, g[f"r{l}"] = self._spec_bwd(dpre, c[f"zh{l}"], P[f"r{l}"])
            dv = dv_w + dv_s
        g["pb"] = dv.sum(axis=(0, 2, 3))
        g["pw"] = np.einsum("bcpq,bipq->ci", dv, x)
        return loss, self.pack(g)
