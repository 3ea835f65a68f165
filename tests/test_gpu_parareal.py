"""Learned parareal on the device (hawassi_c_amd/parareal.py) on a 64 x 64
shoal basin: the exactness of the first k slices after k iterations, the
fallback to the serial fine solver, and what the trained spectral neural
operator gains over the same parareal without it (the depth-blended linear
propagator alone).  The fine solver runs at rtol 1e-4 / atol 1e-5: at the
looser 1e-2 / 1e-3 its adaptive steps amplify a 1e-16 change of the start
state to ~1e-5 after one 5-s slice, so no iteration could agree to 1e-4."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_SLICES = 8


@pytest.fixture(scope="module")
def case():
    from hawassi_c_amd import parareal as PR
    from hawassi_c_amd import workloads as W
    from hawassi_c_amd.fno import FNO
    m = W.shoal(64, shoal_depth=0.7, eps_rel=1e-4, eps_abs=1e-5)
    print("shoal 64: trajectories", flush=True)
    traj, _ = PR.snapshots(m, PR.training_seas(m, range(32)), 30.0)
    base = PR.LinearBlend(m, 5.0)
    X, Y, st = PR.training_pairs(traj, m.depth, 5.0, base)
    # 24 retained modes cover the fine solver's whole anti-alias band at 64^2
    net = FNO(64, 32, 24, 4, 64, max_batch=8, seed=1)
    print(f"training on {X.shape[0]} pairs", flush=True)
    hist, _ = PR.train(net, X, Y, 30, 8, 2e-3)
    assert hist[-1] < 0.2 * hist[0]
    G = PR.CoarseFNO(net, m.depth, st, base)
    u0 = W.shoal_sea(m, 1000, hs=0.04)
    pr = PR.Parareal(m, G, N_SLICES, 5.0)
    ref = pr.fine_serial(u0)
    print("serial reference done", flush=True)
    yield dict(m=m, base=base, G=G, u0=u0, pr=pr, ref=ref)
    pr.close()


def test_first_k_slices_are_exact_after_k_iterations(case):
    pr, ref = case["pr"], case["ref"]
    pr.G = case["base"]
    r = pr.run(case["u0"], tol=0.0, max_iter=2)
    assert r["iterations"] == 2
    # F applied to the same slices as the serial run, up to the rounding of
    # G + F - G and the FP32 transforms' response to it
    for n in (1, 2):
        assert pr.change(ref[n], r["U"][n]) < 1e-5


def test_divergent_coarse_falls_back_to_the_serial_solver(case):
    pr, ref = case["pr"], case["ref"]
    pr.G = lambda eta, phi: (eta, phi)     # the identity: a useless coarse propagator
    r = pr.run(case["u0"], tol=1e-4)
    assert r["fallback"] and not r["converged"]
    assert max(pr.change(ref[n], r["U"][n]) for n in range(1, N_SLICES + 1)) < 1e-5


def test_trained_network_gains_over_the_linear_coarse(case):
    pr, ref = case["pr"], case["ref"]
    out = {}
    for name in ("fno", "linear"):
        pr.G = case["G"] if name == "fno" else case["base"]
        out[name] = pr.run(case["u0"], tol=1e-4, reference=ref)
        print(name, out[name]["iterations"], ["%.1e" % e for e in out[name]["error"]], flush=True)
    f, l = out["fno"], out["linear"]
    assert f["converged"] and not f["fallback"]
    assert f["iterations"] < N_SLICES
    # converged to the serial fine solution
    assert max(pr.change(ref[n], f["U"][n]) for n in range(1, N_SLICES + 1)) <= 1e-4
    # what the network gains: a smaller error at each of the first iterations
    # and no more iterations than the method without it
    assert f["error"][0] < 0.8 * l["error"][0]
    assert f["error"][1] < 0.8 * l["error"][1]
    assert f["iterations"] <= l["iterations"]
