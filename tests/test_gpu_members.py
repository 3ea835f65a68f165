"""Contexts holding several states of one model (hw_ctx_create_2d_members):
every member takes exactly the steps and ends on exactly the state that a
context of its own in graph mode gives (the kernels compute each member with
the same arithmetic; the error norm's maximum does not depend on the block
partition), including members that start at other times or are already at
t1; the single-state entry points refuse such a context."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(n, seeds, order=2):
    from hawassi_c_amd import parareal as PR
    from hawassi_c_amd import workloads as W
    m = W.shoal(n, shoal_depth=0.6, eps_rel=1e-4, eps_abs=1e-5)
    m.order = order
    codec = PR.Codec(m)
    return m, [codec.to_state(*W.shoal_sea(m, s, hs=0.04 + 0.01 * i, theta0=0.4 * i)) for i, s in enumerate(seeds)]


def _single(m, y0, prec, t0, t1, h0):
    from hawassi_c_amd.model2d import Hawassi2D
    d = Hawassi2D(m, prec)
    d.set_step_mode(True)
    d.set_state(y0)
    t, h = d.integrate(t0, t1, h0) if t0 < t1 else (t0, h0)
    y, st = d.get_state(), d.stats()
    d.close()
    return y, st, t, h


@pytest.mark.parametrize("n,prec,order", [(64, 1, 2), (64, 0, 2), (128, 1, 2), (64, 1, 3), (64, 1, 1)])
def test_members_equal_single_contexts(n, prec, order):
    from hawassi_c_amd.model2d import Hawassi2D
    m, ys = _case(n, [3, 4, 5, 6], order)
    t0 = [0.0, 0.3, 0.0, 2.0]          # the last member starts at t1: it must not move
    h0 = [0.05, 0.02, 0.1, 0.05]
    print(f"members n={n} prec={prec} order={order}", flush=True)
    d = Hawassi2D(m, prec, members=4)
    d.set_state(np.stack(ys))
    t, h = d.integrate_members(t0, 2.0, h0)
    yb = d.get_state()
    stb = d.stats()
    d.close()
    acc = rej = 0
    for z in range(4):
        ys_z, st, tz, hz = _single(m, ys[z], prec, t0[z], 2.0, h0[z])
        assert np.array_equal(yb[z], ys_z), z
        assert t[z] == tz and h[z] == hz
        acc += st["steps_accepted"]
        rej += st["steps_rejected"]
    assert np.array_equal(yb[3], ys[3]) or np.allclose(yb[3], ys[3], rtol=0, atol=0)
    assert (stb["steps_accepted"], stb["steps_rejected"]) == (acc, rej)


def test_members_refuse_single_state_calls():
    from hawassi_c_amd.model2d import Hawassi2D
    m, ys = _case(64, [1, 2])
    d = Hawassi2D(m, 1, members=2)
    d.set_state(np.stack(ys))
    with pytest.raises(Exception, match="members"):
        d.rhs(0.0, ys[0])
    with pytest.raises(Exception, match="members"):
        d.integrate(0.0, 1.0, 0.05)
    with pytest.raises(Exception):
        d.set_step_mode(False)
    d.close()


def test_members_refuse_breaking():
    from hawassi_c_amd import workloads as W
    from hawassi_c_amd.model2d import Hawassi2D
    with pytest.raises(Exception, match="breaking"):
        Hawassi2D(W.harbour(128, t_end=2.0), 1, members=2)


def test_member_contexts_of_uneven_sizes_in_one_batch():
    """Realizations on grids of different sizes, one members-context per
    size, advanced together by hw_integrate_batch in graph mode: each member
    ends exactly where a context of its own takes it."""
    from hawassi_c_amd.model2d import Hawassi2D, integrate_batch
    (m1, ys1), (m2, ys2) = _case(64, [3, 4, 5]), _case(128, [6, 7])
    d1 = Hawassi2D(m1, 1, members=3)
    d2 = Hawassi2D(m2, 1, members=2)
    d1.set_state(np.stack(ys1))
    d2.set_state(np.stack(ys2))
    t0 = [0.0, 0.2, 0.0, 0.0, 0.1]
    h0 = [0.05, 0.05, 0.02, 0.05, 0.05]
    print("uneven member contexts 3 x 64^2 + 2 x 128^2", flush=True)
    t, h = integrate_batch([d1, d2], t0, 1.5, h0)
    y1, y2 = d1.get_state(), d2.get_state()
    d1.close()
    d2.close()
    for z, (m, y0, yb) in enumerate([(m1, ys1[0], y1[0]), (m1, ys1[1], y1[1]), (m1, ys1[2], y1[2]),
                                     (m2, ys2[0], y2[0]), (m2, ys2[1], y2[1])]):
        ys_z, _, tz, hz = _single(m, y0, 1, t0[z], 1.5, h0[z])
        assert np.array_equal(yb, ys_z), z
        assert t[z] == tz == 1.5 and h[z] == hz
