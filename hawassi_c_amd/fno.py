"""Host binding of the spectral neural operator (csrc/fno.hip, hw_fno_* in
include/hawassi_hip.h): the learned coarse propagator of the parareal
driver.  No counterpart in the reference.

``FNO(n, ...)`` owns one device network; ``forward``, ``loss_grad``,
``set_dataset`` and ``train_step`` call the device kernels (BF16 matrix-core
products, FP32 accumulation).  Arrays are FP32 numpy ``[batch][ch][n][n]``.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib

_fp = ct.POINTER(ct.c_float)


def _f(a):
    return a.ctypes.data_as(_fp)


class FNO:
    def __init__(self, n: int, width: int = 32, modes: int = 32, layers: int = 4, hidden: int = 128,
                 in_ch: int = 3, out_ch: int = 2, max_batch: int = 8, seed: int = 1):
        self.lib = _lib.load()
        c = _lib.HwFnoConfig()
        c.n, c.width, c.modes, c.layers, c.hidden = n, width, modes, layers, hidden
        c.in_ch, c.out_ch, c.max_batch, c.seed = in_ch, out_ch, max_batch, seed
        h = ct.c_void_p()
        _lib.check(self.lib.hw_fno_create(ct.byref(c), ct.byref(h)), "hw_fno_create")
        self.ctx = h
        self.n, self.in_ch, self.out_ch, self.max_batch = n, in_ch, out_ch, max_batch
        self.modes = modes
        self.nparam = int(self.lib.hw_fno_param_count(self.ctx))

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.hw_fno_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, split: bool):
        """Products of later calls in split BF16 (three MFMA, ~1e-5 relative)
        when `split`, else in BF16 (~4e-3 relative, the training default)."""
        _lib.check(self.lib.hw_fno_set_precision(self.ctx, _lib.HW_FNO_SPLIT if split else _lib.HW_FNO_BF16),
                   "hw_fno_set_precision")

    def gemm_flops(self) -> float:
        """Floating-point operations of the matrix products launched so far."""
        return float(self.lib.hw_fno_gemm_flops(self.ctx))

    def gemm_bytes(self) -> float:
        """Compulsory HBM bytes of the matrix products launched so far."""
        return float(self.lib.hw_fno_gemm_bytes(self.ctx))

    def params(self) -> np.ndarray:
        p = np.empty(self.nparam, dtype=np.float32)
        _lib.check(self.lib.hw_fno_get_params(self.ctx, _f(p)), "hw_fno_get_params")
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        _lib.check(self.lib.hw_fno_set_params(self.ctx, _f(p)), "hw_fno_set_params")

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        b = x.shape[0]
        y = np.empty((b, self.out_ch, self.n, self.n), dtype=np.float32)
        for i in range(0, b, self.max_batch):
            j = min(b, i + self.max_batch)
            xi = np.ascontiguousarray(x[i:j])
            yi = np.empty((j - i, self.out_ch, self.n, self.n), dtype=np.float32)
            _lib.check(self.lib.hw_fno_forward(self.ctx, j - i, _f(xi), _f(yi)), "hw_fno_forward")
            y[i:j] = yi
        return y

    def loss_grad(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.nparam, dtype=np.float32)
        loss = ct.c_double()
        _lib.check(self.lib.hw_fno_loss_grad(self.ctx, x.shape[0], _f(x), _f(t), ct.byref(loss), _f(g)),
                   "hw_fno_loss_grad")
        return loss.value, g

    def set_dataset(self, x: np.ndarray, y: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        _lib.check(self.lib.hw_fno_set_dataset(self.ctx, x.shape[0], _f(x), _f(y)), "hw_fno_set_dataset")

    def train_step(self, idx, lr: float) -> float:
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        loss = ct.c_double()
        _lib.check(self.lib.hw_fno_train_step(self.ctx, idx.size, idx.ctypes.data_as(ct.POINTER(ct.c_int)),
                                              float(lr), ct.byref(loss)), "hw_fno_train_step")
        return loss.value


def gemm_check(a: np.ndarray, b: np.ndarray, split: bool = False) -> np.ndarray:
    """C = A B through the device's MFMA GEMM, BF16 or split BF16 (for its tests)."""
    lib = _lib.load()
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    M, K = a.shape
    N = b.shape[1]
    c = np.empty((M, N), dtype=np.float32)
    _lib.check(lib.hw_fno_gemm_check(M, N, K, _f(a), _f(b), _f(c), _lib.HW_FNO_SPLIT if split else _lib.HW_FNO_BF16),
               "hw_fno_gemm_check")
    return c


def _main():
    """python -m hawassi_c_amd.fno [--n 512] [--steps 20]: Adam steps on
    seeded random fields (batch 8), for profiling the training kernels;
    prints the matrix-product flops and the steps' wall time."""
    import argparse
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--modes", type=int, default=32)
    a = ap.parse_args()
    g = np.random.default_rng(0)
    x = g.standard_normal((a.batch, 3, a.n, a.n)).astype(np.float32)
    y = g.standard_normal((a.batch, 2, a.n, a.n)).astype(np.float32)
    net = FNO(a.n, 32, a.modes, 4, 128, max_batch=a.batch, seed=1)
    net.set_dataset(x, y)
    idx = np.arange(a.batch)
    net.train_step(idx, 1e-3)
    f0, b0 = net.gemm_flops(), net.gemm_bytes()
    w0 = time.perf_counter()
    for i in range(a.steps):
        net.train_step(idx, 1e-3)
    wall = time.perf_counter() - w0
    fl, by = net.gemm_flops() - f0, net.gemm_bytes() - b0
    print(f"fno n={a.n} batch={a.batch}: {a.steps} Adam steps in {wall:.3f} s, {fl / a.steps / 1e9:.2f} GFLOP and "
          f"{by / a.steps / 1e9:.2f} GB (compulsory) of matrix products per step, {fl / wall / 1e12:.2f} TFLOP/s "
          f"over the steps", flush=True)


if __name__ == "__main__":
    _main()
