"""HS1 dynamics (rhs_AB1_2DF_uv / rhs_AB1_2DB_uv, the linear model) on the
device against the FP64 oracle, through the C ABI.  Tolerances as for HS2:
FP64 relative L2 <= 1e-11, FP32 <= 3e-5 per RHS field."""
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

TOL_RHS = {HW_PREC_FP64: 1e-11, HW_PREC_FP32: 3e-5}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def model(kind, n):
    m = getattr(W, kind)(n, t_end=8.0)
    m.order = 1
    return m


@pytest.mark.parametrize("prec", [HW_PREC_FP64, HW_PREC_FP32])
@pytest.mark.parametrize("kind,n", [("short_crested", 64), ("short_crested", 128), ("harbour", 64), ("harbour", 128)])
def test_hs1_rhs_parity(prec, kind, n):
    say("hs1 rhs", kind, n, prec)
    m = model(kind, n)
    y = W.random_state(m, 7, kfrac=0.8)
    ref = HS.oracle_half(O.Oracle2D(m).rhs(1.234, y), n, n)
    dev = Hawassi2D(m, prec)
    got = HS.oracle_half(dev.rhs(1.234, y), n, n)
    dev.close()
    for f in range(3):
        e = relerr(got[f], ref[f])
        say("  field", f, "relL2", e)
        assert e < TOL_RHS[prec]


@pytest.mark.parametrize("kind", ["short_crested", "harbour"])
def test_hs1_trajectory(kind):
    """From rest with the influx on to t = 1, 2 s at eps 1e-10 (FP64)."""
    say("hs1 trajectory", kind)
    n = 64
    m = model(kind, n)
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
        assert t_d == t1 and t_o == t1
    got = HS.oracle_half(dev.get_state(), n, n)
    ref = HS.oracle_half(y, n, n)
    dev.close()
    for f in range(3):
        e = relerr(got[f], ref[f])
        say("  field", f, "relL2", e)
        assert e < 1e-8
