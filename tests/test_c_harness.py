"""The C-ABI used from C as the reference uses its RHS: tests/c/gsl_style_driver.c
includes include/hawassi_hip.h, links libhawassi_hip.so, registers hw_rhs as
the function of a gsl_odeiv2_system-shaped struct (2d/main.c:204-222) and
advances the full host state with a host-side GSL rkf45 / control_y /
evolve_apply loop (ode_solver_uv, 2d/fun_odesolver.c:2565-2647).  The CPU
test compiles and links it (C99, -pedantic -Werror); the GPU tests run it and
compare with the device-resident integrator (hw_integrate) and with the FP64
oracle's own GSL restatement."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "hawassi_c_amd")
SRC = os.path.join(ROOT, "tests", "c", "gsl_style_driver.c")


def build_driver(out_dir):
    if not os.path.exists(os.path.join(LIB_DIR, "libhawassi_hip.so")):
        pytest.skip("libhawassi_hip.so is not built")
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    exe = os.path.join(str(out_dir), "gsl_style_driver")
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
           SRC, "-o", exe, "-L", LIB_DIR, "-lhawassi_hip", "-Wl,-rpath," + LIB_DIR, "-Wl,-rpath-link,/opt/rocm/lib",
           "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def write_model(path, m, y0, t0, t1, h0, precision, full_grid=0):
    n = m.nx * m.ny
    with open(path, "wb") as f:
        f.write(struct.pack("<6i", m.nx, m.ny, int(m.flat), int(getattr(m, "order", 2)), precision, int(full_grid)))
        from hawassi_c_amd.setup2d import GRAV
        f.write(struct.pack("<7d", GRAV, m.cpeak, m.eps_abs, m.eps_rel, t0, t1, h0))
        arrs = [m.kx, m.ky, m.aal, m.fbdy, m.wallchar, m.chi]
        if m.flat:
            arrs += [m.c2d, m.l2d]
        else:
            o = m.ops
            arrs += [o["gam_min"], o["gam_plus"], o["csq_min"], o["csq_plus"], o["om2_min"], o["om2_plus"]]
        for a in arrs:
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        assert y0.size == 6 * n
        f.write(np.ascontiguousarray(y0, dtype="<f8").tobytes())


def read_out(path, dim):
    raw = open(path, "rb").read()
    y = np.frombuffer(raw[:8 * dim], dtype="<f8")
    acc, rej, calls = struct.unpack("<3q", raw[8 * dim:8 * dim + 24])
    return y, acc, rej, calls


def test_driver_compiles_and_links(tmp_path):
    build_driver(tmp_path)


def _case(flat):
    from hawassi_c_amd import parareal as PR
    from hawassi_c_amd import workloads as W
    m = W.shoal(64, shoal_depth=1.0 if flat else 0.6, eps_rel=1e-6, eps_abs=1e-7)
    y0 = PR.Codec(m).to_state(*W.shoal_sea(m, 11, hs=0.06, theta0=0.3))
    return m, y0


@pytest.mark.gpu
@pytest.mark.parametrize("flat", [False, True])
def test_gsl_style_host_loop_matches_the_device_integrator_and_the_oracle(tmp_path, flat):
    from hawassi_c_amd import HW_PREC_FP64
    from hawassi_c_amd.model2d import Hawassi2D
    from oracle.oracle import Oracle2D
    exe = build_driver(tmp_path)
    m, y0 = _case(flat)
    t1, h0 = 1.5, 0.05
    model_file, out_file = str(tmp_path / "model.bin"), str(tmp_path / "out.bin")
    write_model(model_file, m, y0, 0.0, t1, h0, HW_PREC_FP64)
    print(f"C driver, flat={flat}", flush=True)
    r = subprocess.run([exe, model_file, out_file], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout.strip(), flush=True)
    y, acc, rej, calls = read_out(out_file, y0.size)
    # dydt_in once per accepted step, five stages and dydt_out per attempt
    assert calls == acc + 6 * (acc + rej) and acc > 5
    # the device-resident integrator (FP64) from the same state
    dev = Hawassi2D(m, HW_PREC_FP64)
    dev.set_state(y0)
    dev.integrate(0.0, t1, h0)
    yd = dev.get_state()
    st = dev.stats()
    dev.close()
    h = m.nx // 2 + 1
    stored = np.zeros((6, m.ny, m.nx), dtype=bool)
    stored[:, :, :h] = True
    stored = stored.ravel()
    scale = np.abs(yd[stored]).max()
    # same steps, same stored half (the host loop's norm also runs over the
    # redundant half, which hw_rhs fills with the mirror, so it decides alike)
    assert (acc, rej) == (st["steps_accepted"], st["steps_rejected"])
    assert np.abs(y[stored] - yd[stored]).max() <= 1e-12 * scale
    # the oracle's own GSL loop: its RHS fills the redundant half as the
    # reference does (values that differ from the mirror by O(1/n) at low k,
    # the wavenumber grid not being antisymmetric), and its norm runs over
    # them, so its step sequence may differ by a step (flat case: 51 accepted
    # steps against 50); the states then agree to the step-control level
    yo = y0.copy()
    rc, to, _, (oacc, orej) = Oracle2D(m).integrate(0.0, t1, h0, yo, m.eps_abs, m.eps_rel)
    assert rc == 0 and to == t1
    print("oracle steps", oacc, orej, "device", acc, rej, flush=True)
    assert abs(oacc - acc) <= 1
    assert np.abs(y[stored] - yo[stored]).max() <= 1e-8 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("flat", [False, True])
def test_gsl_style_host_loop_on_the_full_grid_takes_the_oracles_steps(tmp_path, flat):
    """With hw_config_2d.full_grid = 1, hw_rhs writes the reference's own
    values in the redundant columns too, so the host GSL loop's norm sees what
    the reference's sees: the C driver takes the oracle's step sequence (the
    flat case included, where the mirrored half makes it take one step less)
    and ends on the oracle's full state, every value."""
    from hawassi_c_amd import HW_PREC_FP64
    from oracle.oracle import Oracle2D
    exe = build_driver(tmp_path)
    m, y0 = _case(flat)
    t1, h0 = 1.5, 0.05
    model_file, out_file = str(tmp_path / "model.bin"), str(tmp_path / "out.bin")
    write_model(model_file, m, y0, 0.0, t1, h0, HW_PREC_FP64, full_grid=1)
    print(f"C driver on the full grid, flat={flat}", flush=True)
    r = subprocess.run([exe, model_file, out_file], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    y, acc, rej, calls = read_out(out_file, y0.size)
    yo = y0.copy()
    rc, to, _, (oacc, orej) = Oracle2D(m).integrate(0.0, t1, h0, yo, m.eps_abs, m.eps_rel)
    print("oracle steps", oacc, orej, "C driver", acc, rej, flush=True)
    assert rc == 0 and to == t1
    assert (acc, rej) == (oacc, orej)
    assert np.abs(y - yo).max() <= 1e-10 * np.abs(yo).max()
