"""The FP64 oracle of the spectral neural operator (oracle/fno_oracle.py), no
GPU: its backward pass against central differences of its forward pass, and
its spectral convolution against the closed form (identity mode weights pass
a field whose spectrum lies in the retained modes through unchanged)."""
import numpy as np
import pytest

from oracle.fno_oracle import FNOOracle


def small(seed=0, n=16, width=6, modes=4, layers=3, hidden=10, batch=2):
    o = FNOOracle(n, width, modes, layers, hidden)
    g = np.random.default_rng(seed)
    p = 0.3 * g.standard_normal(o.nparam)
    x = g.standard_normal((batch, 3, n, n))
    t = g.standard_normal((batch, 2, n, n))
    return o, p, x, t


def test_backward_matches_central_differences():
    o, p, x, t = small()
    loss, grad = o.loss_grad(p, x, t)
    g = np.random.default_rng(1)
    picks = list(g.choice(o.nparam, 60, replace=False))
    for name, off, shp in o.layout:          # one entry of every parameter block, too
        picks.append(off + int(np.prod(shp)) // 2)
    h = 1e-5
    for i in picks:
        e = np.zeros(o.nparam)
        e[i] = h
        fd = (o.loss_grad(p + e, x, t)[0] - o.loss_grad(p - e, x, t)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_backward_check_catches_a_wrong_mode_weight():
    """Plant a mistake (the spectral weights' gradient of the kx > 0 modes
    counted once, not twice) in the backward pass: the central-difference
    check above must fail on such a weight."""
    o, p, x, t = small()
    orig = o._spec_bwd

    def wrong(dy, zh, R):
        dv, dR = orig(dy, zh, R)
        return dv, dR * np.where(np.arange(o.m) == 0, 1.0, 0.5)
    o._spec_bwd = wrong
    _, gw = o.loss_grad(p, x, t)
    o._spec_bwd = orig
    off = dict((nm, (of, shp)) for nm, of, shp in o.layout)["r0"][0]
    i = off + 2                                # part 0, in 0, out 0, ky index 0, kx 2
    h = 1e-5
    e = np.zeros(o.nparam)
    e[i] = h
    fd = (o.loss_grad(p + e, x, t)[0] - o.loss_grad(p - e, x, t)[0]) / (2 * h)
    assert abs(fd - gw[i]) > 0.3 * abs(fd)


def test_identity_mode_weights_pass_a_band_limited_field():
    n, m, C = 32, 6, 3
    o = FNOOracle(n, C, m, 1, 4)
    g = np.random.default_rng(3)
    spec = np.zeros((2, C, n, n), dtype=complex)
    keep_y = o.kys
    for ky in keep_y:
        for kx in range(m):
            spec[:, :, ky, kx] = g.standard_normal((2, C)) + 1j * g.standard_normal((2, C))
    spec[:, :, n - m, 0] = 0.0                 # its mirror (ky = m, kx = 0) is not a retained mode
    v = np.real(np.fft.ifft2(spec))           # real field; its spectrum: the kept modes and their mirrors
    R = np.zeros((2, C, C, 2 * m, m))
    for i in range(C):
        R[0, i, i] = 1.0
    _, _, y = o._spec_fwd(v, R)
    assert np.abs(y - v).max() < 1e-12 * np.abs(v).max()
    # the same field through a truncation one mode narrower loses energy
    o2 = FNOOracle(n, C, m - 1, 1, 4)
    _, _, y2 = o2._spec_fwd(v, R[:, :, :, np.r_[0:m - 1, m + 1:2 * m], :m - 1])
    assert np.abs(y2 - v).max() > 1e-3 * np.abs(v).max()
