"""Device step control replayed from a HIP graph (hw_set_step_mode
HW_STEP_GRAPH) against the host step loop (HW_STEP_HOST, gsl evolve_apply
restated on the host): the same accepted and rejected steps and the same
state, with and without the influx source, in one context, in a batch, and
when graph and host segments alternate on one context; breaking is refused.
Both paths run the same rkf45 arithmetic on the device (the stage and
norm kernels are compiled without floating-point contraction, so the
host-stepped and graph-stepped kernels give the same bits); the step-size
update uses pow() on the device in one and on the host in the other, so a
step size may differ in its last bit.  The error estimate is a small
difference of the stage derivatives, so such a difference reaches the later
step sizes amplified: in the FP64 oracle on this 64^2 shoal, one step size
changed by one ulp moves the final suggested step by up to 8.6e-9
(relative) and the state by 6e-14 after 2 s, hence 2e-8 on the step size
and 1e-12 on the state in FP64.  The FP32 transforms turn a difference that
reaches the stage inputs into rounding-level changes (a 1e-16 change of the
state moves an FP32 run by ~1e-7 after a few hundred evaluations), hence the
FP32 tolerance of 2e-6 on the state, and the same 2e-8 on the step size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _shoal(n, prec_fp32=True, seed=5):
    from hawassi_c_amd import parareal as PR
    from hawassi_c_amd import workloads as W
    m = W.shoal(n, shoal_depth=0.6, eps_rel=1e-4, eps_abs=1e-5)
    return m, PR.Codec(m).to_state(*W.shoal_sea(m, seed, hs=0.05, theta0=0.5))


def _run(m, y0, graph, t1, prec, h0=0.05, segments=None):
    from hawassi_c_amd.model2d import Hawassi2D
    d = Hawassi2D(m, prec)
    d.set_state(y0)
    t, h = 0.0, h0
    for (t_end, g) in (segments or [(t1, graph)]):
        d.set_step_mode(g)
        t, h = d.integrate(t, t_end, h)
    y = d.get_state()
    st = d.stats()
    d.close()
    return y, st, t, h


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("n,prec", [(64, 0), (64, 1), (256, 1)])
def test_graph_steps_match_the_host_loop(n, prec):
    m, y0 = _shoal(n)
    print(f"shoal {n} prec {prec}", flush=True)
    yh, sh, th, hh = _run(m, y0, False, 2.0, prec)
    yg, sg, tg, hg = _run(m, y0, True, 2.0, prec)
    assert (sg["steps_accepted"], sg["steps_rejected"]) == (sh["steps_accepted"], sh["steps_rejected"])
    assert tg == th == 2.0
    assert abs(hg - hh) <= 2e-8 * abs(hh)
    assert _rel(yg, yh) < (1e-12 if prec == 0 else 2e-6)
    # K0 in every attempt: one evaluation more per rejected step
    assert sg["rhs_evals"] == 6 * (sg["steps_accepted"] + sg["steps_rejected"])


def test_graph_steps_with_the_influx_source():
    from hawassi_c_amd import workloads as W
    m = W.short_crested(128, t_end=4.0)
    y0 = np.zeros(6 * 128 * 128)
    print("short crested 128", flush=True)
    yh, sh, _, _ = _run(m, y0, False, 3.0, 1, h0=0.1)
    yg, sg, _, _ = _run(m, y0, True, 3.0, 1, h0=0.1)
    assert np.abs(yh).max() > 0
    assert (sg["steps_accepted"], sg["steps_rejected"]) == (sh["steps_accepted"], sh["steps_rejected"])
    assert _rel(yg, yh) < 2e-6


def test_graph_and_host_segments_alternate():
    m, y0 = _shoal(64)
    yh, sh, _, hh = _run(m, y0, False, 3.0, 1)
    ya, sa, ta, ha = _run(m, y0, None, None, 1, segments=[(1.0, True), (1.7, False), (2.4, True), (3.0, False)])
    yb, sb, tb, hb = _run(m, y0, None, None, 1, segments=[(1.0, False), (1.7, True), (2.4, False), (3.0, True)])
    # segment ends clip steps, so compare the two alternations with a host run cut the same way
    yc, sc, _, _ = _run(m, y0, None, None, 1, segments=[(1.0, False), (1.7, False), (2.4, False), (3.0, False)])
    for y, s in ((ya, sa), (yb, sb)):
        assert s["steps_accepted"] == sc["steps_accepted"] and s["steps_rejected"] == sc["steps_rejected"]
        assert _rel(y, yc) < 2e-6
    assert ta == tb == 3.0


def test_graph_batch_equals_single_contexts():
    from hawassi_c_amd.model2d import Hawassi2D, integrate_batch
    from hawassi_c_amd import HW_PREC_FP32
    cases = [_shoal(64, seed=s) for s in (1, 2, 3)]
    singles = [_run(m, y0, True, 2.5, 1) for m, y0 in cases]
    ds = []
    for m, y0 in cases:
        d = Hawassi2D(m, HW_PREC_FP32)
        d.set_state(y0)
        d.set_step_mode(True)
        ds.append(d)
    t, h = integrate_batch(ds, [0.0, 0.0, 0.0], 2.5, [0.05, 0.05, 0.05])
    for d, (ys, ss, ts, hs), tb, hb in zip(ds, singles, t, h):
        assert np.array_equal(d.get_state(), ys)
        st = d.stats()
        assert (st["steps_accepted"], st["steps_rejected"]) == (ss["steps_accepted"], ss["steps_rejected"])
        assert tb == ts and hb == hs
        d.close()


def test_graph_mode_refuses_breaking():
    from hawassi_c_amd import workloads as W
    from hawassi_c_amd.model2d import Hawassi2D
    from hawassi_c_amd import HW_PREC_FP32
    m = W.harbour(128, t_end=2.0)
    d = Hawassi2D(m, HW_PREC_FP32)
    d.set_state(np.zeros(6 * 128 * 128))
    d.set_step_mode(True)
    with pytest.raises(Exception, match="breaking"):
        d.integrate(0.0, 1.0, 0.05)
    d.close()
