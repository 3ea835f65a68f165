"""Host side of the learned parareal (hawassi_c_amd/parareal.py), no GPU:
the (eta, phi) <-> device-state codec, and the linear coarse propagators
against the FP64 oracle's right-hand side (oracle/oracle.py, the restatement
of rhs_AB2_2DF_uv) on the flat periodic basin, where linear propagation is
exact: the propagator's time derivative at dt = 0, taken by central
differences in dt, must equal the oracle's RHS of a small-amplitude sea."""
import numpy as np
import pytest

from hawassi_c_amd import parareal as PR
from hawassi_c_amd import workloads as W
from oracle.oracle import Oracle2D


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("shoal_depth", [0.7, 1.0])
def test_codec_round_trip(shoal_depth):
    m = W.shoal(64, shoal_depth=shoal_depth)
    c = PR.Codec(m)
    e, p = W.shoal_sea(m, 4, hs=0.05, theta0=0.7)
    e2, p2 = c.from_state(c.to_state(e, p))
    assert rel(e2, e) < 1e-13 and rel(p2, p) < 1e-13


def _flat_sea(n=64, amp=1e-6, seed=2):
    m = W.shoal(n, shoal_depth=1.0)
    e, p = W.shoal_sea(m, seed, hs=4 * amp, theta0=0.4)
    return m, e, p


def test_linear_flat_derivative_matches_the_oracle_rhs():
    m, e, p = _flat_sea()
    c = PR.Codec(m)
    dy = Oracle2D(m).rhs(0.0, c.to_state(e, p))
    de_o, dp_o = c.from_state(dy)
    h = 1e-3
    fwd, bwd = PR.LinearFlat(m, h)(e, p), PR.LinearFlat(m, -h)(e, p)
    de, dp = (fwd[0] - bwd[0]) / (2 * h), (fwd[1] - bwd[1]) / (2 * h)
    # central difference O(h^2 w^3) and the oracle's quadratic terms at
    # amplitude 1e-6: both far below 1e-5
    assert rel(de, de_o) < 1e-5 and rel(dp, dp_o) < 1e-5


def test_linear_flat_check_catches_a_wrong_depth():
    """Planted mistake: the flat propagator at 0.8 of the basin's depth
    changes the derivative of eta far beyond the tolerance above."""
    m, e, p = _flat_sea()
    c = PR.Codec(m)
    de_o, _ = c.from_state(Oracle2D(m).rhs(0.0, c.to_state(e, p)))
    h = 1e-3
    fwd, bwd = PR.LinearFlat(m, h, d0=0.8)(e, p), PR.LinearFlat(m, -h, d0=0.8)(e, p)
    assert rel((fwd[0] - bwd[0]) / (2 * h), de_o) > 1e-2


def test_linear_propagators_on_the_flat_basin():
    """Group property G(2 dt) = G(dt)^2 of the exact linear map (to 1e-9:
    the reference wavenumber grid is not symmetric in ky, |k(-ky)| differs
    from |k(ky)| by one part in n, and the real inverse keeps the Hermitian
    part of column kx = 0 as the device inverse does); the depth blend
    reduces to the flat propagator."""
    m, e, p = _flat_sea(amp=0.01)
    g1, g2 = PR.LinearFlat(m, 2.5), PR.LinearFlat(m, 5.0)
    a = g1(*g1(e, p))
    b = g2(e, p)
    assert rel(a[0], b[0]) < 1e-9 and rel(a[1], b[1]) < 1e-9
    assert rel(g1(e, p)[0], b[0]) > 0.1
    blend = PR.LinearBlend(m, 5.0)
    c = blend(e, p)
    assert rel(c[0], b[0]) < 1e-12 and rel(c[1], b[1]) < 1e-12


def test_linear_blend_weights_partition_unity():
    m = W.shoal(64, shoal_depth=0.4)
    b = PR.LinearBlend(m, 5.0, ndepth=5)
    s = np.sum(b.wts, axis=0)
    assert np.allclose(s, 1.0, atol=1e-14)
    assert min(w.min() for w in b.wts) >= 0.0
    # the deepest reference depth carries the far field alone
    assert np.allclose(b.wts[-1][0, 0], 1.0, atol=1e-6)
