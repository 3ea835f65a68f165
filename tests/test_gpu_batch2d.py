"""hw_integrate_batch: several 2-D contexts advanced in lockstep give exactly
what hw_integrate gives for each alone (same kernels, same order per
context, each with its own step control), for mixed sizes and operators."""
import sys

import numpy as np
import pytest

from hawassi_c_amd import HW_PREC_FP32, HW_PREC_FP64
from hawassi_c_amd import workloads as W
from hawassi_c_amd._lib import HwError
from hawassi_c_amd.model2d import Hawassi2D, integrate_batch

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _models():
    # flat open sea (rhs_AB2_2DF_uv) realizations with different seeds, plus a
    # two-depth harbour with walls and breaking (rhs_AB2_2DB_uv) of another size
    ms = [W.short_crested(128, seed=s, t_end=4.0) for s in (3, 4, 5)]
    ms.append(W.harbour(64, t_end=4.0))
    return ms


@pytest.mark.parametrize("prec", [HW_PREC_FP64, HW_PREC_FP32])
def test_batch_equals_individual(prec):
    say("batch2d", prec)
    ms = _models()
    ys = [W.random_state(m, 40 + i, amp=0.02) for i, m in enumerate(ms)]
    alone, stats_alone = [], []
    for m, y in zip(ms, ys):
        d = Hawassi2D(m, prec)
        d.set_state(y)
        t, h = 0.0, 0.05
        for t1 in (0.5, 1.0):
            t, h = d.integrate(t, t1, h)
        alone.append((t, h, d.get_state()))
        stats_alone.append(d.stats())
        d.close()
    ds = [Hawassi2D(m, prec) for m in ms]
    for d, y in zip(ds, ys):
        d.set_state(y)
    t, h = [0.0] * len(ds), [0.05] * len(ds)
    for t1 in (0.5, 1.0):
        t, h = integrate_batch(ds, t, t1, h)
    for i, d in enumerate(ds):
        ta, ha, ya = alone[i]
        st = d.stats()
        say("  ctx", i, "t", t[i], "h", h[i], "steps", st["steps_accepted"], st["steps_rejected"])
        assert t[i] == ta and h[i] == ha
        assert np.array_equal(d.get_state(), ya)
        for k in ("rhs_evals", "steps_accepted", "steps_rejected"):
            assert st[k] == stats_alone[i][k]
    for d in ds:
        d.close()


def test_batch_short_crested_1024_x8():
    """The open-sea workload at its full size, eight realizations in one
    batch: all reach t1 and stay finite."""
    say("batch2d 1024 x8")
    ms = [W.short_crested(1024, seed=3 + b, t_end=2.0) for b in range(8)]
    ds = [Hawassi2D(m, HW_PREC_FP32) for m in ms]
    for d in ds:
        d.set_state(np.zeros(6 * 1024 * 1024))
    t, h = integrate_batch(ds, [0.0] * 8, 1.0, [0.1] * 8)
    assert all(x == 1.0 for x in t)
    for d in ds:
        eta, _, _ = d.fields()
        assert np.isfinite(eta).all() and np.abs(eta).max() > 0
        d.close()


def test_batch_empty_and_errors():
    assert integrate_batch([], [], 1.0, []) == ([], [])
    m = W.short_crested(64, t_end=2.0)
    d = Hawassi2D(m, HW_PREC_FP64)
    d.set_state(np.zeros(6 * 64 * 64))
    with pytest.raises(HwError):
        integrate_batch([d], [0.0], 1.0, [-0.1])   # step direction opposite to the interval
    d.close()
