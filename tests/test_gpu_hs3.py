"""HS3 dynamics (rhs_AB3_2DF_uv / rhs_AB3_2DB_uv) on the device against the
FP64 oracle, through the C ABI.

Besides the whole right-hand side (FP64 relative L2 <= 1e-11, FP32 <= 3e-5
per field, as for HS2) the HS3 increment R(HS3) - R(HS2) is compared on its
own, so that a wrong third-order term cannot hide under the AB2 terms:
FP64 <= 1e-7 of the increment's norm, FP32 <= 5e-2 on a state whose
increment is ~3e-4 of the right-hand side."""
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
TOL_INC = {HW_PREC_FP64: 1e-7, HW_PREC_FP32: 5e-2}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rhs_pair(m, prec, t, y, n):
    out = {}
    for order in (2, 3):
        m.order = order
        dev = Hawassi2D(m, prec)
        out[order] = HS.oracle_half(dev.rhs(t, y), n, n)
        dev.close()
    return out


def ref_pair(m, t, y, n):
    out = {}
    for order in (2, 3):
        m.order = order
        out[order] = HS.oracle_half(O.Oracle2D(m).rhs(t, y), n, n)
    return out


@pytest.mark.parametrize("prec", [HW_PREC_FP64, HW_PREC_FP32])
@pytest.mark.parametrize("kind,n", [("short_crested", 64), ("short_crested", 128), ("harbour", 64), ("harbour", 128)])
def test_hs3_rhs_parity(prec, kind, n):
    say("hs3 rhs", kind, n, prec)
    m = getattr(W, kind)(n, t_end=8.0)
    if prec == HW_PREC_FP32:
        y = W.random_state(m, 7, amp=2.0, kfrac=0.15)
    else:
        y = W.random_state(m, 7, kfrac=0.8)
    ref = ref_pair(m, 1.234, y, n)
    got = rhs_pair(m, prec, 1.234, y, n)
    for f in range(3):
        e = relerr(got[3][f], ref[3][f])
        inc = relerr(got[3][f] - got[2][f], ref[3][f] - ref[2][f])
        say("  field", f, "relL2", e, "increment relL2", inc,
            "increment size", relerr(ref[3][f], ref[2][f]))
        assert e < TOL_RHS[prec]
        assert inc < TOL_INC[prec]


def test_hs3_breaking_sequence_fp64():
    """The breaking flux sees delphiH with the HS3 terms (the reference sums
    deluH3 into delphiH_hat before fun_sbreaking)."""
    n = 128
    say("hs3 breaking sequence", n)
    m = W.harbour(n, t_end=8.0, kbc=0.2)
    m.order = 3
    y = W.random_state(m, 7, amp=2.0, kfrac=0.15)
    o = O.Oracle2D(m)
    d = Hawassi2D(m, HW_PREC_FP64)
    o.set_tprev(0.0)
    d.set_tprev(0.0)
    active = 0
    for t in (1.0, 1.05, 1.1, 1.3):
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
    d.close()
    assert active > 0, "the state did not trigger breaking; the test would be vacuous"


@pytest.mark.parametrize("kind", ["short_crested", "harbour"])
def test_hs3_trajectory(kind):
    """From rest with the influx on to t = 1, 2 s at eps 1e-10 (FP64)."""
    say("hs3 trajectory", kind)
    n = 64
    m = getattr(W, kind)(n, t_end=8.0)
    m.order = 3
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


@pytest.mark.parametrize("prec", [HW_PREC_FP32])
def test_hs3_2048_increment_is_cubic(prec):
    """At 2048^2 (the size the throughput is quoted on) the HS3 increment is
    cubic in the state amplitude: R3(2y) - R2(2y) = 8 (R3(y) - R2(y)) with the
    source off, to FP32 accuracy."""
    n = 2048
    say("hs3 2048 cubic", prec)
    m = W.harbour(n, t_end=8.0, breaking=False)
    m.source_on = False
    y = W.random_state(m, 3, amp=1.0, kfrac=0.1)
    inc = []
    for s in (1.0, 2.0):
        r = rhs_pair(m, prec, 0.5, s * y, n)
        inc.append([r[3][f] - r[2][f] for f in range(3)])
    for f in range(3):
        e = relerr(inc[1][f], 8.0 * inc[0][f])
        say("  field", f, "cubic relL2", e)
        assert e < 5e-2
