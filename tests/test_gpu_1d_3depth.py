"""The 1-D path with three reference depths (Up3IP and the middle term of
HSS, 1d/functions.c:2247-2340, 2929-2975) on the device against the FP64
oracle: right-hand sides (FP64 <= 1e-11, FP32 <= 2e-5, and the middle
depth's share alone <= 1e-8 in FP64), a mixed batch of two- and three-depth
realizations, and a trajectory with identical step counts."""
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


def model3(n, breaking=False, t_end=4.0):
    m = W1.slope(n, t_end=t_end, breaking=breaking)
    m.ops = dict(S.up3ip(m.depth, m.k, 2 * math.pi / 2.0))
    return m


def two(m):
    return S.Model1D(**{**m.__dict__, "ops": {k: v for k, v in m.ops.items() if k not in ("gam_mid", "csq_mid")}})


@pytest.mark.parametrize("n,prec", [(256, HW_PREC_FP64), (2048, HW_PREC_FP64), (8192, HW_PREC_FP64),
                                    (1024, HW_PREC_FP32), (16384, HW_PREC_FP32)])
def test_rhs1d_three_depth_parity(n, prec):
    say("rhs1d 3depth", n, prec)
    m = model3(n)
    y = W1.random_state(m, 3, amp=0.05, kfrac=0.3)
    ref3 = O.Oracle1D(m).rhs(0.37, y)
    ref2 = O.Oracle1D(two(m)).rhs(0.37, y)
    d = Hawassi1D([m, two(m)], prec)
    got3 = d.rhs(0.37, y, 0)
    got2 = d.rhs(0.37, y, 1)
    e = relerr(got3, ref3)
    inc = relerr(got3 - got2, ref3 - ref2)
    say("  relL2", e, "middle-depth share relL2", inc, "share size", relerr(ref3, ref2))
    assert e < (1e-11 if prec == HW_PREC_FP64 else 2e-5)
    if prec == HW_PREC_FP64:
        assert inc < 1e-8


def test_trajectory1d_three_depth_with_breaking():
    say("trajectory1d 3depth")
    m = model3(512, breaking=True)
    d = Hawassi1D([m], HW_PREC_FP64, t0=0.0, h0=1e-4)
    o = O.Oracle1D(m)
    y = np.zeros(4 * m.n)
    t, h = 0.0, 1e-4
    for t1 in np.arange(0.2, 1.01, 0.2):
        d.integrate(t1)
        rc, t, h, _ = o.integrate(t, t1, h, y)
        assert rc == 0
    got = d.get_state(0)
    e = relerr(got, y)
    st = d.stats(0)
    say("  relL2", e, "device rhs", st["rhs_evals"], "oracle rhs", o.rhs_count())
    assert e < 1e-9 and st["rhs_evals"] == o.rhs_count()
