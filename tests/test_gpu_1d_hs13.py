"""The 1-D rhs_HsF1 and rhs_HsF3 (1d/rhs.c:1-94, 243-399) on the device
against the FP64 oracle: right-hand sides (FP64 <= 1e-11, FP32 <= 2e-5; the
HS3 increment R(HS3) - R(HS2) alone <= 1e-8 in FP64), two- and three-depth
bottoms, a batch mixing the three dynamics, and trajectories with identical
step counts (HS3 with breaking)."""
import math

import numpy as np
import pytest
from conftest import relerr

from hawassi_c_amd import HW_PREC_FP32, HW_PREC_FP64
from hawassi_c_amd import setup1d as S
from hawassi_c_amd import workloads1d as W1
from hawassi_c_amd.model1d import Hawassi1D
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def model(kind, n, order, mid=False):
    m = getattr(W1, kind)(n, t_end=4.0, **({} if kind == "channel" else {"ndepth": 3 if mid else 2}))
    m.order = order
    return m


def with_order(m, order):
    return S.Model1D(**{**m.__dict__, "order": order})


@pytest.mark.parametrize("kind,n,prec,mid", [
    ("channel", 256, HW_PREC_FP64, False), ("channel", 2048, HW_PREC_FP64, False),
    ("slope", 512, HW_PREC_FP64, False), ("slope", 4096, HW_PREC_FP64, False), ("slope", 1024, HW_PREC_FP64, True),
    ("channel", 1024, HW_PREC_FP32, False), ("slope", 16384, HW_PREC_FP32, False), ("slope", 2048, HW_PREC_FP32, True),
])
def test_rhs1d_hs1_hs3_parity(kind, n, prec, mid):
    say("rhs1d hs1/hs3", kind, n, prec, "mid", mid)
    m2 = model(kind, n, 2, mid)
    y = W1.random_state(m2, 3, amp=0.1, kfrac=0.3)
    ms = [with_order(m2, o) for o in (1, 2, 3)]
    d = Hawassi1D(ms, prec)
    got = [d.rhs(0.37, y, r) for r in range(3)]
    ref = [O.Oracle1D(m).rhs(0.37, y) for m in ms]
    tol = 1e-11 if prec == HW_PREC_FP64 else 2e-5
    for o, g_, r_ in zip((1, 2, 3), got, ref):
        e = relerr(g_, r_)
        say("  order", o, "relL2", e)
        assert e < tol
    inc = relerr(got[2] - got[1], ref[2] - ref[1])
    say("  HS3 increment relL2", inc, "size", relerr(ref[2], ref[1]))
    if prec == HW_PREC_FP64:
        assert inc < 1e-8


@pytest.mark.parametrize("order", [1, 3])
def test_trajectory1d_hs1_hs3_fp64(order):
    say("trajectory1d order", order)
    ms = [with_order(W1.slope(512, t_end=4.0, breaking=True), order), with_order(W1.channel(256, t_end=4.0), order)]
    d = Hawassi1D(ms, HW_PREC_FP64, t0=0.0, h0=1e-4)
    os_ = [O.Oracle1D(m) for m in ms]
    ys = [np.zeros(4 * m.n) for m in ms]
    ts, hs = [0.0, 0.0], [1e-4, 1e-4]
    for t1 in np.arange(0.2, 1.21, 0.2):
        d.integrate(t1)
        for r in range(2):
            rc, ts[r], hs[r], _ = os_[r].integrate(ts[r], t1, hs[r], ys[r])
            assert rc == 0
    for r in range(2):
        e = relerr(d.get_state(r), ys[r])
        st = d.stats(r)
        say("  r", r, "relL2", e, "rhs", st["rhs_evals"], os_[r].rhs_count(), "steps", st)
        assert e < 1e-9
        # without breaking anywhere in the batch the device skips the
        # stepper's dydt_out evaluation (side-effect free, its value unused):
        # one RHS fewer per attempt than the oracle, which makes every call
        attempts = st["steps_accepted"] + st["steps_rejected"]
        expect = os_[r].rhs_count() - (attempts if order == 1 else 0)
        assert st["rhs_evals"] == expect
