"""Three reference depths (depth_ref = 3: funOprt_DispInterpolation_3p and
HSSgen_3depth) on the device against the oracle, for the HS1, HS2 and HS3
right-hand sides.  Besides the whole right-hand side (FP64 <= 1e-11, FP32
<= 3e-5) the middle depth's share R(3 depths) - R(2 depths) is compared on
its own (FP64 <= 1e-8 of its norm)."""
import math
import sys

import numpy as np
import pytest
from conftest import relerr

import halfspec as HS
from hawassi_c_amd import HW_PREC_FP32, HW_PREC_FP64
from hawassi_c_amd import setup2d as S
from hawassi_c_amd import workloads as W
from hawassi_c_amd.model2d import Hawassi2D
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MID = ("gam_mid", "csq_mid", "om2_mid")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def model(n, order, breaking=False):
    m = W.harbour(n, t_end=8.0, breaking=breaking, kbc=0.2)
    m.order = order
    o = S.disp_interpolation_3p(m.depth, m.kx, m.ky, 2 * math.pi / 8.0)
    m.ops = {**m.ops, **{k: o[k] for k in ("gam_min", "gam_plus", "csq_min", "csq_plus", "om2_min", "om2_plus") + MID}}
    return m


def two_depth(m):
    return {k: v for k, v in m.ops.items() if k not in MID}


@pytest.mark.parametrize("prec", [HW_PREC_FP64, HW_PREC_FP32])
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("n", [64, 128])
def test_three_depth_rhs_parity(prec, order, n):
    say("3depth rhs", n, "order", order, "prec", prec)
    m = model(n, order)
    y = W.random_state(m, 7, amp=0.5, kfrac=0.5)
    ops3 = m.ops
    ref3 = HS.oracle_half(O.Oracle2D(m).rhs(1.234, y), n, n)
    d = Hawassi2D(m, prec)
    got3 = HS.oracle_half(d.rhs(1.234, y), n, n)
    d.close()
    m.ops = two_depth(m)
    ref2 = HS.oracle_half(O.Oracle2D(m).rhs(1.234, y), n, n)
    d = Hawassi2D(m, prec)
    got2 = HS.oracle_half(d.rhs(1.234, y), n, n)
    d.close()
    m.ops = ops3
    for f in range(3):
        e = relerr(got3[f], ref3[f])
        inc = relerr(got3[f] - got2[f], ref3[f] - ref2[f])
        say("  field", f, "relL2", e, "middle-depth share relL2", inc, "share size", relerr(ref3[f], ref2[f]))
        assert e < (1e-11 if prec == HW_PREC_FP64 else 3e-5)
        if prec == HW_PREC_FP64:
            assert inc < 1e-8


def test_three_depth_breaking_sequence_fp64():
    n = 128
    say("3depth breaking sequence", n)
    m = model(n, 2, breaking=True)
    y = W.random_state(m, 7, amp=2.0, kfrac=0.15)
    o = O.Oracle2D(m)
    d = Hawassi2D(m, HW_PREC_FP64)
    o.set_tprev(0.0)
    d.set_tprev(0.0)
    active = 0
    for t in (1.0, 1.05, 1.1):
        ref = o.rhs(t, y)
        got = d.rhs(t, y)
        Bo = o.B()
        Bd, fd = d.breaking_state()
        assert fd == o.flagbr() and np.abs(Bd - Bo).max() < 1e-12
        active += int((Bo > 0).sum() > 0)
        e = [relerr(a, b) for a, b in zip(HS.oracle_half(got, n, n), HS.oracle_half(ref, n, n))]
        say("   t", t, "relL2", e)
        assert max(e) < 1e-10
    d.close()
    assert active > 0


@pytest.mark.parametrize("order", [2, 3])
def test_three_depth_trajectory(order):
    n = 64
    say("3depth trajectory", order)
    m = model(n, order)
    m.eps_abs = m.eps_rel = 1e-10
    o = O.Oracle2D(m)
    y = np.zeros(6 * n * n)
    dev = Hawassi2D(m, HW_PREC_FP64)
    dev.set_state(y)
    t_o, h_o, t_d, h_d = 0.0, 0.05, 0.0, 0.05
    for t1 in (1.0, 2.0):
        rc, t_o, h_o, _ = o.integrate(t_o, t1, h_o, y, m.eps_abs, m.eps_rel)
        assert rc == 0
        t_d, h_d = dev.integrate(t_d, t1, h_d)
    got = HS.oracle_half(dev.get_state(), n, n)
    ref = HS.oracle_half(y, n, n)
    dev.close()
    e = [relerr(got[f], ref[f]) for f in range(3)]
    say("  relL2", e)
    assert max(e) < 1e-8
