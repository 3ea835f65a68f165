"""hw_fft2_forward_noalloc: the zero-allocation forward transform on device
arrays (scratch in the library's static device array), checked against the
allocating entry point hw_fft2_forward on two calls in a row, the second on a
smaller grid, in FP64 and FP32; sizes beyond the scratch are refused.

The test's own device buffers (the caller's inputs and outputs) come from the
HIP runtime the library links (libamdhip64 through ctypes), so the process
holds one HIP runtime only."""
import ctypes as ct

import numpy as np
import pytest

from hawassi_c_amd import _lib

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2     # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    def __init__(self):
        _lib.load()                       # loads libamdhip64.so.7 as the library's dependency
        self.rt = ct.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        self.rt.hipFree.argtypes = [ct.c_void_p]
        self.rt.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = ct.c_void_p()
        assert self.rt.hipMalloc(ct.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, _H2D) == 0
        return p

    def get(self, p, shape):
        out = np.empty(shape)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, _D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _host_forward(lib, nx, ny, prec, phys):
    re, im = np.empty((ny, nx)), np.empty((ny, nx))
    _lib.check(lib.hw_fft2_forward(nx, ny, prec, _lib.ptr(np.ascontiguousarray(phys)), _lib.ptr(re), _lib.ptr(im)),
               "hw_fft2_forward")
    return re + 1j * im


@pytest.mark.parametrize("prec,tol", [(_lib.HW_PREC_FP64, 1e-13), (_lib.HW_PREC_FP32, 2e-6)])
def test_noalloc_forward_matches_the_allocating_entry_point(prec, tol):
    lib = _lib.load()
    hip = Hip()
    g = np.random.default_rng(11)
    try:
        for (nx, ny) in ((512, 512), (128, 64)):          # the second call on a smaller grid
            print(f"noalloc forward {nx}x{ny} prec {prec}", flush=True)
            phys = g.standard_normal((ny, nx))
            ref = _host_forward(lib, nx, ny, prec, phys)
            d_p = hip.put(phys)
            d_re = hip.put(np.full((ny, nx), np.nan))
            d_im = hip.put(np.full((ny, nx), np.nan))
            _lib.check(lib.hw_fft2_forward_noalloc(nx, ny, prec, d_p, d_re, d_im, None), "hw_fft2_forward_noalloc")
            assert hip.rt.hipDeviceSynchronize() == 0
            got = hip.get(d_re, (ny, nx)) + 1j * hip.get(d_im, (ny, nx))
            assert np.isfinite(got).all()
            err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            assert err < tol, (nx, ny, err)
            # and against numpy's FFT (the reference's forward is the unnormalised exp(-i) DFT)
            assert np.linalg.norm(got - np.fft.fft2(phys)) / np.linalg.norm(ref) < (1e-12 if prec == 0 else 2e-6)
    finally:
        hip.free()


def test_noalloc_forward_refuses_grids_beyond_its_scratch():
    lib = _lib.load()
    hip = Hip()
    try:
        d = hip.alloc(8 * 1024 * 1024)
        rc = lib.hw_fft2_forward_noalloc(1024, 1024, _lib.HW_PREC_FP32, d, d, d, None)
        assert rc == 1
        assert b"scratch" in lib.hw_last_error()
    finally:
        hip.free()
