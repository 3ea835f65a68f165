"""Breaking (2d/breaking.c, fun_sbreaking) on the device vs the oracle.

The same steep state is evaluated at increasing times with tprev fixed
below them, so each call runs breaking_process and the crest state machine
advances (detection, tracking, termination); B, flagbr and the RHS are
compared after every call.  FP64 device path: RHS relL2 <= 1e-10, B to
1e-12.  Gated calls (tprev >= t) must reuse the stale Sb.
"""
import sys

import numpy as np
import pytest
from conftest import relerr

import halfspec as HS
from hawassi_c_amd import HW_PREC_FP32, HW_PREC_FP64
from hawassi_c_amd import workloads as W
from hawassi_c_amd.model2d import Hawassi2D
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n,kbc", [(128, 0.2), (256, 0.3)])
def test_breaking_sequence_fp64(n, kbc):
    say("breaking_sequence", n, kbc)
    m = W.harbour(n, t_end=8.0, kbc=kbc)
    y = W.random_state(m, 7, amp=2.0, kfrac=0.15)
    o = O.Oracle2D(m)
    d = Hawassi2D(m, HW_PREC_FP64)
    o.set_tprev(0.0)
    d.set_tprev(0.0)
    active = 0
    for t in (1.0, 1.05, 1.1, 1.3, 2.0, 3.5):
        ref = o.rhs(t, y)
        got = d.rhs(t, y)
        Bo = o.B()
        Bd, fd = d.breaking_state()
        say(f"  t={t} flagbr oracle={o.flagbr()} device={fd} B>0 {int((Bo > 0).sum())}/{int((Bd > 0).sum())}")
        assert fd == o.flagbr()
        assert np.abs(Bd - Bo).max() < 1e-12
        active += int((Bo > 0).sum() > 0)
        e = [relerr(a, b) for a, b in zip(HS.oracle_half(got, n, n), HS.oracle_half(ref, n, n))]
        say("   relL2", e)
        assert max(e) < 1e-10
    assert active > 0, "the state did not trigger breaking; the test would be vacuous"
    # gated call: stale Sb is reused, nothing recomputed
    o.set_tprev(10.0)
    d.set_tprev(10.0)
    ref = o.rhs(4.0, y)
    got = d.rhs(4.0, y)
    assert max(relerr(a, b) for a, b in zip(HS.oracle_half(got, n, n), HS.oracle_half(ref, n, n))) < 1e-10


def test_breaking_sequence_fp32():
    say("breaking_sequence_fp32")
    n = 256
    m = W.harbour(n, t_end=8.0, kbc=0.3)
    y = W.random_state(m, 7, amp=2.0, kfrac=0.15)
    o = O.Oracle2D(m)
    d = Hawassi2D(m, HW_PREC_FP32)
    o.set_tprev(0.0)
    d.set_tprev(0.0)
    for t in (1.0, 1.05, 1.1):
        ref = o.rhs(t, y)
        got = d.rhs(t, y)
        Bo = o.B()
        Bd, fd = d.breaking_state()
        # the FP32 path may flip a node at a threshold; the fields must agree overall
        assert relerr(Bd, Bo) < 1e-3
        e = [relerr(a, b) for a, b in zip(HS.oracle_half(got, n, n), HS.oracle_half(ref, n, n))]
        say("   relL2", e, "B relL2", relerr(Bd, Bo))
        assert max(e) < 1e-4


def test_breaking_trajectory_fp64():
    """Adaptive integration with breaking on: tprev follows the driver (the
    first stage of each step is gated, the later stages and rkf45's dydt_out
    evaluation run breaking_process).  Both sides integrate to the same t1."""
    say("breaking_trajectory")
    n = 128
    m = W.harbour(n, t_end=8.0, kbc=0.2)
    m.eps_abs = m.eps_rel = 1e-6
    y0 = W.random_state(m, 7, amp=2.0, kfrac=0.15)
    o = O.Oracle2D(m)
    d = Hawassi2D(m, HW_PREC_FP64)
    y = y0.copy()
    d.set_state(y0)
    o.set_tprev(0.0)
    d.set_tprev(0.0)
    rc, t_o, h_o, _ = o.integrate(0.0, 0.004, 0.001, y, m.eps_abs, m.eps_rel)
    t_d, h_d = d.integrate(0.0, 0.004, 0.001)
    assert rc == 0 and t_o == t_d == 0.004
    Bd, fd = d.breaking_state()
    say("  flagbr", o.flagbr(), fd, "B>0", int((o.B() > 0).sum()), int((Bd > 0).sum()),
        "oracle breaking calls", o.break_calls())
    assert o.break_calls() > 0 and o.flagbr() == 2 and fd == 2
    e = [relerr(a, b) for a, b in zip(HS.oracle_half(d.get_state(), n, n), HS.oracle_half(y, n, n))]
    say("  relL2", e)
    assert max(e) < 1e-5
