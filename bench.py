"""Benchmark: 2-D variable-depth harbour, 2048 x 2048, AB2 (HS2) right-hand side
with two-reference-depth operators, advanced by the device-resident RKF45
(FP32 transforms/operators, FP64 state, stage sums and error norm).

One "step" is one accepted RKF45 step (gsl_odeiv2_evolve_apply) of the whole
grid.  Prints one JSON line (rank 0).  Headline value: simulated seconds per
wall second; RHS evaluations per second are reported beside it.

Roofline: the dominant unit is one RHS evaluation (the 9-kernel launch
sequence), timed with HIP events on the library's stream around every
fourth evaluation of the timed region (an event record idles the GPU for a
few microseconds, so the others run without).  Its algorithmic HBM bytes per evaluation are
B_rhs = T_alg (4N + 8H) + C_phys 4N + C_spec 4H + 2 S (input state read and
derivative write), T_alg = 36, C_phys = 10, C_spec = 5, S = 3*16*H,
N = nx*ny, H = (nx/2+1)*ny.  With --order 3 (the AB3 "HS3" right-hand side,
13 kernels) T_alg = 47 (LLphi and dgp inverted; A, g-A, g+A, eta^2 M0u w,
eta^2 M0v w and dE3 forward; the HSS tail of A: two inverses, one forward)
and C_phys = 14 (physical M0u, M0v written once and read once).  The whole-step figure (adding the RKF45 vector
traffic, 5.5 S per evaluation) is reported as roofline_job.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# HBM bytes actually moved per RHS evaluation of the 2048^2 harbour, FP32 path:
# FETCH_SIZE and WRITE_SIZE counter passes (rocprofv3 --pmc) over this script
# (FETCH_SIZE scaled by 2.0, WRITE_SIZE by 1.0, the factors that make a known
# 512 MiB copy read right), summed over the evaluation's kernels: the nine
# transform passes (2046.7 MB) and the peak search's share (29.1 MB).
RHS_TRAFFIC_BYTES_2048 = 2075.8e6
# the same for the AB3 (HS3) right-hand side
RHS_TRAFFIC_BYTES_2048_HS3 = 2963.3e6
# HBM bytes per right-hand side of the default 1-D batch (256 sloping-beach
# realizations of seed 7, two reference depths, every realization active):
# the three size classes' largest dispatches in the FETCH_SIZE (x2.0) and
# WRITE_SIZE (x1.0) counter passes
RHS_TRAFFIC_BYTES_SLOPE1D = 1255.8e6
METRIC = "sim-sec/wall-sec + RHS evals/s, 2D 2048^2 FP32 vs FP64 oracle at relL2<=1e-4"


def algorithmic_bytes(nx: int, ny: int, t_alg: int = 36, c_phys: int = 10, c_spec: int = 5):
    N = nx * ny
    H = (nx // 2 + 1) * ny
    S = 3 * 16 * H
    transforms = t_alg * (4 * N + 8 * H)
    coeff = c_phys * 4 * N + c_spec * 4 * H
    rhs_only = transforms + coeff + 2 * S
    job = rhs_only + 5.5 * S
    return rhs_only, job


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--spinup", type=float, default=150.0, help="simulated seconds integrated from rest before warmup")
    ap.add_argument("--acc-window", type=float, default=60.0,
                    help="harbour2048: FP32 vs FP64 over this many simulated seconds ending at the spin-up time")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", default="harbour2048",
                    choices=["harbour2048", "slope1d", "opensea1024x8", "parareal512", "ensemble", "refcase2d",
                             "kin1d"],
                    help="harbour2048: the 2-D headline; slope1d: the 1-D batch of uneven realizations; "
                         "opensea1024x8: eight flat-bottom 1024^2 realizations advanced in lockstep; "
                         "parareal512: learned parareal on the 512^2 shoal basin (trains the network first); "
                         "ensemble: 84 sea states on 128^2/256^2/512^2 grids, members and HIP-graph steps; "
                         "refcase2d: the reference case 2dFBr001_Irreg (600x300, breaking) whole, full-grid path; "
                         "kin1d: the 1-D internal kinematics at one output time (4096 nodes, 90 levels)")
    ap.add_argument("--pr-seeds", type=int, default=33, help="parareal512: training trajectories (61 snapshots each)")
    ap.add_argument("--pr-epochs", type=int, default=30, help="parareal512: training epochs")
    ap.add_argument("--pr-window", type=int, default=8, help="parareal512: slices per parareal window")
    ap.add_argument("--pr-shoal-depth", type=float, default=0.4, help="parareal512: depth on the shoal's crest")
    ap.add_argument("--nreal", type=int, default=256, help="slope1d: realizations in the batch")
    ap.add_argument("--depths", type=int, default=2, choices=[2, 3], help="slope1d: reference depths")
    ap.add_argument("--round-limit", type=int, default=0,
                    help="slope1d: rounds per output interval before the run stops with an error (0: none)")
    ap.add_argument("--order", type=int, default=2, choices=[2, 3],
                    help="harbour2048: dynamics, 2 = HS2 (rhs_AB2_2DB_uv), 3 = HS3 (rhs_AB3_2DB_uv)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="launch one elementwise copy kernel (torch.neg) over this many MiB of fp32 first, so a "
                         "counter pass over this command holds a dispatch of known HBM traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard over independent sea-state realizations, one context per GPU)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_mib:
        import torch
        x = torch.ones(args.copy_probe_mib * (1 << 18), device="cuda")
        y = torch.empty_like(x)
        torch.neg(x, out=y)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written")
        del x, y
    if args.workload == "slope1d":
        return bench_slope1d(args)
    if args.workload == "opensea1024x8":
        return bench_opensea(args)
    if args.workload == "parareal512":
        return bench_parareal(args)
    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "refcase2d":
        return bench_refcase2d(args)
    if args.workload == "kin1d":
        return bench_kin1d(args)

    from hawassi_c_amd import HW_PREC_FP32, HW_PREC_FP64
    from hawassi_c_amd import workloads as W
    from hawassi_c_amd.model2d import Hawassi2D

    n = args.n
    spin = args.spinup
    t_acc0 = max(0.0, spin - args.acc_window)
    t_end = spin + 40.0
    log(f"setup harbour {n}x{n}")
    model = W.harbour(n, t_end=t_end)
    model.order = args.order
    dev = Hawassi2D(model, HW_PREC_FP32)
    dim = 6 * n * n
    dev.set_state(np.zeros(dim))
    t, h = 0.0, 0.05
    # spin-up from rest: the influx fills the domain, the waves pass the
    # breakwater gap (x = 1000 m, reached at ~80 s) and run onto the shallows;
    # the state at t_acc0 starts the accuracy comparison
    w0 = time.perf_counter()
    state_acc0, h_acc0 = None, None
    for tc in np.arange(10.0, spin + 1e-9, 10.0):
        if state_acc0 is None and tc > t_acc0 + 1e-9:
            t, h = dev.integrate(t, t_acc0, h)
            state_acc0, h_acc0 = dev.get_state(), h
        t, h = dev.integrate(t, float(tc), h)
        st = dev.stats()
        log(f"spin-up t={t:.1f} h={h:.4f} rhs={st['rhs_evals']} rejected={st['steps_rejected']} "
            f"wall {time.perf_counter() - w0:.1f} s")
    if state_acc0 is None:
        state_acc0, h_acc0 = dev.get_state(), h
    for _ in range(args.warmup):
        t, h = dev.evolve_apply(t, t_end, h)
    dev.sync()
    dev.reset_stats()
    t_start, h_start = t, h
    state_start = dev.get_state()
    _, flag_start = dev.breaking_state()
    dev.sync()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        t, h = dev.evolve_apply(t, t_end, h)
    dev.sync()
    wall = time.perf_counter() - w0
    st = dev.stats()
    sim = t - t_start
    log(f"timed {args.steps} steps: wall {wall:.3f} s, sim {sim:.3f} s, {st}")
    rhs_per_s = st["rhs_evals"] / wall
    avg_rhs_ms = st["rhs_ms_total"] / max(st.get("rhs_timed") or st["rhs_evals"], 1)
    b_rhs, b_job = algorithmic_bytes(n, n) if args.order == 2 else algorithmic_bytes(n, n, t_alg=47, c_phys=14)
    achieved = b_rhs / (avg_rhs_ms * 1e-3) / 1e9
    job = rhs_per_s * b_job / 1e9
    out = {
        "metric": METRIC,
        "value": sim / wall,
        "unit": "sim-s/wall-s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"harbour {n}x{n}: 1.5 km square, depth 20->4 m + seeded bumps, two-depth "
                               + ("AB2 (HS2)" if args.order == 2 else "AB3 (HS3)") + " operator, breakwater R=0.5 with 100 m gap, directional JONSWAP line influx; "
                               "RKF45 eps_rel 1e-2 eps_abs 1e-3 (reference tolerances); timed after a spin-up from rest "
                               "with the sea developed past the breakwater",
                   "nx": n, "ny": n, "state": "fp64 half spectrum", "sim_window_s": [t_start, t],
                   "breaking_flag_at_start": flag_start},
        "rhs_evals_per_s": rhs_per_s,
        "rhs_evals": st["rhs_evals"],
        "steps_rejected": st["steps_rejected"],
        "avg_rhs_ms": avg_rhs_ms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (RHS_TRAFFIC_BYTES_2048 if args.order == 2 else RHS_TRAFFIC_BYTES_2048_HS3)
                                if n == 2048 else None,
                     "unit_of_work": "one RHS evaluation (%d kernels), %.1f MB algorithmic"
                                     % (9 if args.order == 2 else 13, b_rhs / 1e6)},
        "roofline_job": {"achieved": job, "frac": job / HBM_PEAK_GBS, "unit": "GB/s",
                         "bytes_per_rhs": b_job},
    }
    dev.close()
    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    if not args.no_accuracy:
        # FP32 vs FP64 on the device over the developed-sea window t_acc0..spin,
        # both from the same FP32-run state (fresh breaking histories)
        acc = {"window_s": [t_acc0, spin], "reference": "FP64 device path (oracle-checked in tests)"}
        res = {}
        for name, prec in (("fp64", HW_PREC_FP64), ("fp32", HW_PREC_FP32)):
            d = Hawassi2D(model, prec)
            d.set_state(state_acc0)
            d.set_tprev(t_acc0)
            w1 = time.perf_counter()
            ta, ha = t_acc0, h_acc0
            for tc in np.arange(t_acc0 + 10.0, spin + 1e-9, 10.0):
                ta, ha = d.integrate(ta, float(tc), ha)
            log(f"accuracy run {name}: t={ta:.2f} wall {time.perf_counter() - w1:.1f} s {d.stats()}")
            res[name] = (physical_fields(model, d.get_state()), d.breaking_state())
            d.close()
        (f64, (B64, fl64)), (f32, (B32, fl32)) = res["fp64"], res["fp32"]
        for k in ("eta", "u", "v", "phi"):
            acc["relL2_" + k] = rel(f32[k], f64[k])
        acc["breaking_flag"] = [int(fl32), int(fl64)]
        acc["relL2_B"] = rel(B32, B64) if np.any(B64) else ("B = 0 in both" if not np.any(B32) else "B64 = 0, B32 != 0")
        acc["pass_1e-4"] = bool(max(acc["relL2_" + k] for k in ("eta", "u", "v", "phi")) <= 1e-4)
        out["accuracy"] = acc

    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from oracle import oracle as O
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        O.set_threads(threads)
        log(f"cpu baseline: oracle, {threads} threads, 3 RKF45 steps from t={t_start:.3f}")
        orc = O.Oracle2D(model)
        orc.set_tprev(t_start)   # the driver's tprev after the previous accepted step
        y = state_start.copy()
        c0 = orc.rhs_count()
        w0 = time.perf_counter()
        tt, hh = t_start, h_start
        y_one = None
        for k in range(3):
            rc, tt, hh, stc = orc.evolve_apply(tt, t_end, hh, y, model.eps_abs, model.eps_rel)
            if k == 0:
                y_one, t_one = y.copy(), tt
            log(f"  oracle step {k + 1}: t={tt:.4f} wall {time.perf_counter() - w0:.1f} s")
        cw = time.perf_counter() - w0
        nr = orc.rhs_count() - c0
        # scaling of the oracle's right-hand side with host threads (one evaluation each)
        scale = {}
        for th in (1, 2, 4, 8):
            O.set_threads(th)
            o2 = O.Oracle2D(model)
            o2.set_tprev(t_start)
            w1 = time.perf_counter()
            o2.rhs(t_start, state_start)
            scale[str(th)] = 1.0 / (time.perf_counter() - w1)
            log(f"  oracle RHS with {th} threads: {1.0 / scale[str(th)]:.2f} s")
        O.set_threads(threads)
        out["cpu_baseline"] = {"value": (tt - t_start) / cw, "unit": "sim-s/wall-s", "cores": threads,
                               "kind": "port",
                               "sample": f"3 RKF45 steps ({nr} RHS incl. GSL dydt_out, {tt - t_start:.4g} s simulated) of "
                                         f"the same {n}x{n} harbour from the same state at t={t_start:.3f} s; FP64 "
                                         f"oracle (restated CPU path, the project's radix-2 FFT -- no FFTW -- with "
                                         f"OpenMP), {cpu_model()}",
                               "rhs_evals_per_s": nr / cw, "wall_s": cw,
                               "rhs_per_s_by_threads": scale}
        # the device from the same state, one RKF45 step, against the oracle's
        # first step: the step the oracle accepted (after any rejected
        # attempts, whose retry size follows from each side's own error norm)
        # is given to the device as its interval, so both end at t_one
        ver = {"t": [t_start, t_one]}
        fo = physical_fields(model, y_one)
        for name, prec in (("fp32", HW_PREC_FP32), ("fp64", HW_PREC_FP64)):
            d = Hawassi2D(model, prec)
            d.set_state(state_start)
            d.set_tprev(t_start)
            td, _ = d.evolve_apply(t_start, t_one, t_one - t_start)
            fd = physical_fields(model, d.get_state())
            ver[name] = {k: rel(fd[k], fo[k]) for k in ("eta", "u", "v", "phi")}
            ver[name]["t_end"] = td
            d.close()
        out["accuracy_vs_oracle_one_step"] = ver
    print(json.dumps(out))


def physical_fields(model, y):
    """eta, u, v and the curl-free potential phi (phi^ = -i (kx u^ + ky v^)/|k|^2)
    of a full-grid state, on the host (the reference's projected inverse)."""
    nx, ny = model.nx, model.ny
    N = nx * ny
    pl = [(y[2 * f * N:(2 * f + 1) * N] + 1j * y[(2 * f + 1) * N:(2 * f + 2) * N]).reshape(ny, nx)[:, : nx // 2 + 1]
          for f in range(3)]
    KX, KY = model.kx[None, : nx // 2 + 1], model.ky[:, None]
    k2 = KX * KX + KY * KY
    ph = -1j * (KX * pl[1] + KY * pl[2]) / np.where(k2 > 1e-12, k2, 1.0)
    ph[0, 0] = 0.0
    out = {}
    for name, s in zip(("eta", "u", "v", "phi"), (pl[0], pl[1], pl[2], ph)):
        s = s.copy()
        s[ny // 2 + 1:, 0] = np.conj(s[1: ny // 2][::-1, 0])   # the projection of column 0
        s[0, 0] = s[0, 0].real
        out[name] = np.fft.irfft2(s, s=(ny, nx))
    return out


def bench_opensea(args):
    """2-D batch: eight realizations (seeds 3..10) of the 2 km flat open sea
    at 1024^2 (rhs_AB2_2DF_uv, directional JONSWAP line influx, absorbing
    strips), FP32 transforms / FP64 state, advanced together by
    hw_integrate_batch through output times every 0.5 s.  A step is one
    output interval of the batch; value = realizations x simulated seconds
    per wall second."""
    from hawassi_c_amd import HW_PREC_FP32
    from hawassi_c_amd import workloads as W
    from hawassi_c_amd.model2d import Hawassi2D, integrate_batch

    nb, n, dt_out = 8, 1024, 0.5
    spin = args.spinup if args.spinup != 150.0 else 10.0
    t_end = spin + (args.warmup + args.steps + 2) * dt_out
    log(f"setup {nb} open-sea realizations {n}x{n}")
    ds = [Hawassi2D(W.short_crested(n, seed=3 + b, t_end=t_end), HW_PREC_FP32) for b in range(nb)]
    for d in ds:
        d.set_state(np.zeros(6 * n * n))
    t, h = [0.0] * nb, [0.1] * nb
    t, h = integrate_batch(ds, t, spin, h)
    log(f"spin-up done t={t[0]:.2f}")
    tc = t[0]
    for _ in range(args.warmup):
        tc += dt_out
        t, h = integrate_batch(ds, t, tc, h)
    for d in ds:
        d.reset_stats()
    t_start = tc
    w0 = time.perf_counter()
    for i in range(args.steps):
        tc += dt_out
        t, h = integrate_batch(ds, t, tc, h)
        log(f"  timed {i + 1}/{args.steps} wall {time.perf_counter() - w0:.2f} s")
    for d in ds:
        d.sync()
    wall = time.perf_counter() - w0
    sts = [d.stats() for d in ds]
    nrhs = sum(st["rhs_evals"] for st in sts)
    sim = tc - t_start
    b_rhs, _ = algorithmic_bytes(n, n, t_alg=16, c_phys=4, c_spec=3)
    out = {
        "metric": "sim-sec/wall-sec (2-D batch, realizations x simulated seconds)",
        "value": nb * sim / wall,
        "unit": "sim-s/wall-s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"open sea batch of {nb}: 2 km square {n}x{n}, depth 30 m, absorbing strips 200 m, "
                               "vertical influx line at x = 250 m, JONSWAP Hs 3 Tp 10 cos^2s s=10, flat AB2 (HS2) "
                               "operator, RKF45 eps_rel 1e-2 eps_abs 1e-3", "sim_window_s": [t_start, tc]},
        "rhs_evals_per_s": nrhs / wall,
        "roofline": {"bound": "hbm", "achieved": nrhs * b_rhs / wall / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": nrhs * b_rhs / wall / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "unit_of_work": "one realization-RHS (5 kernels), %.1f MB algorithmic, whole-job rate" % (b_rhs / 1e6)},
    }
    for d in ds:
        d.close()
    print(json.dumps(out))


def bench_ensemble(args):
    """Batching: random-phase sea states on grids of uneven size advanced
    together.  Realizations of the shoal basin (two reference depths, no
    influx) at 128^2 (64 of them), 256^2 (16) and 512^2 (4): one context per
    size holding that size's realizations as members (every kernel of a
    right-hand side covers all members in one launch), step control on the
    device, each attempted step one HIP-graph replay, the three contexts
    advanced in lockstep by hw_integrate_batch.  A step is one 0.5-s output
    interval of the whole ensemble; value = realizations x simulated seconds
    per wall second.  Beside it: the same realizations one context at a time
    with the host step loop (a sample of each size, scaled by its count)."""
    from hawassi_c_amd import HW_PREC_FP32
    from hawassi_c_amd import parareal as PR
    from hawassi_c_amd import workloads as W
    from hawassi_c_amd.model2d import Hawassi2D, integrate_batch

    sizes = ((128, 64), (256, 16), (512, 4))
    dt_out = 0.5
    ctxs, states, models = [], [], {}
    for n, nb in sizes:
        m = W.shoal(n, shoal_depth=0.6, eps_rel=1e-4, eps_abs=1e-5)
        codec = PR.Codec(m)
        ys = [codec.to_state(*W.shoal_sea(m, 100 * n + i, hs=0.03 + 0.03 * (i % 5) / 4, theta0=0.37 * i))
              for i in range(nb)]
        d = Hawassi2D(m, HW_PREC_FP32, members=nb)
        d.set_state(np.stack(ys))
        ctxs.append(d)
        states.append(ys)
        models[n] = m
        log(f"setup {nb} realizations {n}x{n}")
    nmem = sum(nb for _, nb in sizes)
    t = [0.0] * nmem
    h = [0.02] * nmem
    tc = 0.0
    for _ in range(args.warmup):
        tc += dt_out
        t, h = integrate_batch(ctxs, t, tc, h)
    for d in ctxs:
        d.reset_stats()
    t_start = tc
    w0 = time.perf_counter()
    for i in range(args.steps):
        tc += dt_out
        t, h = integrate_batch(ctxs, t, tc, h)
        log(f"  timed {i + 1}/{args.steps} wall {time.perf_counter() - w0:.2f} s")
    for d in ctxs:
        d.sync()
    wall = time.perf_counter() - w0
    sim = tc - t_start
    rhs = {n: d.stats()["rhs_evals"] for (n, _), d in zip(sizes, ctxs)}
    bytes_alg = sum(rhs[n] * algorithmic_bytes(n, n)[0] for n, _ in sizes)
    # one context at a time, host step loop: a sample per size over the same window
    serial = {}
    for (n, nb), ys in zip(sizes, states):
        k = min(nb, 2)
        timed = 0.0
        for y0 in ys[:k]:
            d = Hawassi2D(models[n], HW_PREC_FP32)
            d.set_state(y0)
            tt, hh = d.integrate(0.0, t_start, 0.02)
            d.sync()
            w2 = time.perf_counter()
            d.integrate(tt, tc, hh)     # the ensemble's timed window only
            d.sync()
            timed += time.perf_counter() - w2
            d.close()
        serial[n] = timed / k * nb
        log(f"  one at a time {n}x{n}: {serial[n]:.2f} s for {nb} realizations (sample of {k})")
    serial_wall = sum(serial.values())
    out = {
        "metric": "sim-sec/wall-sec (2-D ensemble, realizations x simulated seconds)",
        "value": nmem * sim / wall,
        "unit": "sim-s/wall-s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "ensemble of %d shoal-basin sea states on uneven grids: %s; two reference depths, "
                               "RKF45 eps_rel 1e-4 eps_abs 1e-5, members per size in one context, device step "
                               "control from HIP graphs" % (nmem, ", ".join(f"{nb} x {n}^2" for n, nb in sizes)),
                   "sim_window_s": [t_start, tc]},
        "rhs_evals_per_s": sum(rhs.values()) / wall,
        "one_at_a_time_host_loop_s": serial_wall,
        "speedup_vs_one_at_a_time": serial_wall / wall,
        "roofline": {"bound": "hbm", "achieved": bytes_alg / wall / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": bytes_alg / wall / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "unit_of_work": "one realization-RHS at its size (algorithmic bytes), whole-job rate"},
    }
    for d in ctxs:
        d.close()
    print(json.dumps(out))


def bench_refcase2d(args):
    """The reference's own 2-D case 2dFBr001_Irreg (600 x 300, flat bottom,
    HS2, irregular influx from its influx.dat, breaking on), set up from its
    files and run whole (0 -> 150 s) on the full-grid path, FP64 and FP32;
    outputs every 5 s.  Accuracy FP32 against FP64 for eta, phi and the
    breaking mask B at the output times where breaking is active, and one
    second from the FP64 state at the first of them against the FP64 oracle
    (the reference's algorithm on the CPU)."""
    import importlib.util
    import tempfile

    from hawassi_c_amd import HW_PREC_FP32, HW_PREC_FP64
    from hawassi_c_amd import case2d as C
    from hawassi_c_amd.model2d import Hawassi2D
    from oracle import oracle as O

    name = "2dFBr001_Irreg"
    here = os.path.dirname(os.path.abspath(__file__))
    spec = importlib.util.spec_from_file_location("mf2d", os.path.join(here, "tests", "golden", "case2d", "make_fixtures.py"))
    mf = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mf)
    m = C.model_from_input(mf.materialize(name, os.path.join(tempfile.mkdtemp(), name)))
    t_end = float(m.meta["t_end"])
    outs = np.arange(5.0, t_end + 1e-9, 5.0)
    runs = {}
    for prec, tag in ((HW_PREC_FP64, "fp64"), (HW_PREC_FP32, "fp32")):
        d = Hawassi2D(m, prec)
        d.set_state(np.zeros(6 * m.nx * m.ny))
        d.set_tprev(0.0)
        t, h = 0.0, 0.1
        snaps = {}
        w0 = time.perf_counter()
        for t1 in outs:
            t, h = d.integrate(t, float(t1), h)
            B, f = d.breaking_state()
            snaps[float(t1)] = (d.get_state(), B, f, h)
            log(f"  {tag} t={t:.0f} h={h:.4f} flagbr={f} B>0={int((B > 0).sum())}")
        d.sync()
        wall = time.perf_counter() - w0
        runs[tag] = (snaps, wall, d.stats())
        d.close()
    def rel(a, b):
        n = np.linalg.norm(b)
        return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a - b))
    acc = []
    for tt in outs:
        y64, B64, f64, _ = runs["fp64"][0][float(tt)]
        y32, B32, f32, _ = runs["fp32"][0][float(tt)]
        if not ((B64 > 0).any() or (B32 > 0).any()):
            continue
        p64, p32 = physical_fields(m, y64), physical_fields(m, y32)
        acc.append({"t": float(tt), "relL2_eta": rel(p32["eta"], p64["eta"]), "relL2_phi": rel(p32["phi"], p64["phi"]),
                    "relL2_B": rel(B32, B64), "breaking_nodes": [int((B64 > 0).sum()), int((B32 > 0).sum())],
                    "flagbr": [f64, f32]})
    # cpu_baseline leg: the FP64 oracle (restated CPU path) over one second
    # from the FP64 state at the first output time with breaking active,
    # timed, and the device's FP64 run over the same second checked against it
    vs_oracle = None
    if acc and not args.no_cpu_baseline:
        t0 = acc[0]["t"]
        y0, _, _, h0 = runs["fp64"][0][t0]
        o = O.Oracle2D(m)
        d = Hawassi2D(m, HW_PREC_FP64)
        d.set_state(y0)
        o.set_tprev(t0)
        d.set_tprev(t0)
        yo = y0.copy()
        w0 = time.perf_counter()
        rc, to, _, so = o.integrate(t0, t0 + 1.0, h0, yo, m.eps_abs, m.eps_rel)
        wo = time.perf_counter() - w0
        d.integrate(t0, t0 + 1.0, h0)
        yd = d.get_state()
        st = d.stats()
        pd, po = physical_fields(m, yd), physical_fields(m, yo)
        Bd, fd = d.breaking_state()
        d.close()
        vs_oracle = {"window_s": [t0, t0 + 1.0], "steps_oracle": list(so),
                     "steps_device": [st["steps_accepted"], st["steps_rejected"]],
                     "relL2_eta": rel(pd["eta"], po["eta"]), "relL2_phi": rel(pd["phi"], po["phi"]),
                     "relL2_B": rel(Bd, o.B()), "flagbr": [fd, o.flagbr()], "oracle_breaking_calls": o.break_calls(),
                     "oracle_wall_s": wo}
        log(f"  vs oracle: {vs_oracle}")
    s64, w64, st64 = runs["fp64"]
    s32, w32, st32 = runs["fp32"]
    out = {
        "metric": "sim-sec/wall-sec (reference case 2dFBr001_Irreg, whole run, full-grid path)",
        "value": t_end / w32,
        "unit": "sim-s/wall-s",
        "n_gpus": 1, "steps": len(outs), "warmup": 0,
        "ms_per_step": w32 / len(outs) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "the reference's own case files (input.txt, influx.dat)",
        "config": {"workload": "2dFBr001_Irreg 600x300, flat, HS2, breaking (KBC from its input.txt), 0-%g s, "
                               "outputs every 5 s, full-grid path" % t_end},
        "fp64": {"sim_s_per_wall_s": t_end / w64, "wall_s": w64, "rhs_evals": st64["rhs_evals"],
                 "steps": [st64["steps_accepted"], st64["steps_rejected"]]},
        "fp32": {"sim_s_per_wall_s": t_end / w32, "wall_s": w32, "rhs_evals": st32["rhs_evals"],
                 "steps": [st32["steps_accepted"], st32["steps_rejected"]]},
        "accuracy_fp32_vs_fp64_breaking_active": acc,
    }
    if vs_oracle:
        out["cpu_baseline"] = {"value": 1.0 / vs_oracle["oracle_wall_s"], "unit": "sim-s/wall-s",
                               "cores": O.lib().orc_num_threads(), "kind": "port",
                               "sample": "1 s of the case from the device FP64 state at t = %g s (breaking active); "
                                         "FP64 oracle, OpenMP, its own 2/3/5 FFT" % vs_oracle["window_s"][0]}
        out["fp64_vs_oracle"] = vs_oracle
    print(json.dumps(out))


def bench_parareal(args):
    """Learned parareal on the 512^2 shoal basin (workloads.shoal: periodic
    100 m square, depth 1.0 m over a Gaussian shoal, two reference depths,
    no influx, breaking off; fine solver FP32 transforms / FP64 state at
    rtol 1e-4 / atol 1e-5).  The fine solver generates the training
    trajectories (pr_seeds seas of random direction and height, 60 s, a
    snapshot every second), the spectral neural operator (4 Fourier layers,
    32 modes, width 32, GELU, in [eta, phi, depth]) learns the 5-s step
    over the depth-blended linear propagator on the matrix cores, and
    parareal advances one sea over 64 slices of 5 s: in one 64-slice window,
    and in windows of pr_window slices, with and without the network.
    value = the wall-clock speedup of the learned parareal over the serial
    fine solver at the same accuracy (relative L2 <= 1e-4 against it)."""
    from hawassi_c_amd import parareal as PR
    from hawassi_c_amd import workloads as W
    from hawassi_c_amd.fno import FNO

    n, nslices, dtc = 512, 64, 5.0
    m = W.shoal(n, shoal_depth=args.pr_shoal_depth, eps_rel=1e-4, eps_abs=1e-5)
    log(f"shoal basin {n}x{n}, crest depth {args.pr_shoal_depth} m: {args.pr_seeds} training trajectories")
    w0 = time.perf_counter()
    calls = [0]

    def every_tenth(msg):
        calls[0] += 1
        if calls[0] % 10 == 0:
            log(msg)
    traj, _ = PR.snapshots(m, PR.training_seas(m, range(args.pr_seeds)), 60.0, log=every_tenth)
    t_snap = time.perf_counter() - w0
    nsnap = traj.shape[0] * traj.shape[1]
    log(f"{nsnap} snapshots in {t_snap:.1f} s")
    base = PR.LinearBlend(m, dtc)
    w0 = time.perf_counter()
    X, Y, st = PR.training_pairs(traj, m.depth, dtc, base)
    del traj
    log(f"{X.shape[0]} training pairs in {time.perf_counter() - w0:.1f} s")
    net = FNO(n, 32, 32, 4, 128, max_batch=8, seed=1)
    f0 = net.gemm_flops()
    w0 = time.perf_counter()
    hist, val = PR.train(net, X, Y, args.pr_epochs, 8, 2e-3, holdout=56, log=log)
    t_train = time.perf_counter() - w0
    train_flops = net.gemm_flops() - f0
    steps = args.pr_epochs * ((X.shape[0] - 56) // 8)
    del X, Y
    log(f"trained in {t_train:.1f} s: {train_flops / t_train / 1e12:.2f} TFLOP/s of matrix products over the steps")
    G = PR.CoarseFNO(net, m.depth, st, base)
    u0 = W.shoal_sea(m, 1000, hs=0.04)
    pr = PR.Parareal(m, G, nslices, dtc)
    w0 = time.perf_counter()
    ref = pr.fine_serial(u0)
    t_serial = time.perf_counter() - w0
    log(f"serial fine solver, {nslices} slices: {t_serial:.2f} s")
    runs = {}
    for name, g in (("fno", G), ("linear", base)):
        pr.G = g
        for W_ in (nslices, args.pr_window):
            log(f"{name} parareal, windows of {W_} slices")
            r = pr.run_windows(u0, W_, tol=1e-4, reference=ref, log=log)
            key = f"{name}_w{W_}"
            runs[key] = {"window": W_, "iterations": r["iterations"], "iterations_per_window": r["iterations_per_window"],
                         "fallback_windows": r["fallback_windows"], "wall_s": r["wall_s"], "coarse_s": r["coarse_s"],
                         "fine_s": r["fine_s"], "final_rel_l2_vs_serial": r["final_error"],
                         "first_window_error_per_iteration": r["error"][0],
                         "speedup_vs_serial": t_serial / r["wall_s"]}
            log(f"{key}: iterations {r['iterations_per_window']} fallbacks {r['fallback_windows']} wall {r['wall_s']:.2f} s "
                f"(serial {t_serial:.2f}) error {r['final_error']:.2e} first window {['%.1e' % e for e in r['error'][0]]}")
    pr.close()
    best = runs[f"fno_w{args.pr_window}"]
    out = {
        "metric": "learned-parareal wall-clock speedup over the serial fine solver (relative L2 <= 1e-4 against it)",
        "value": best["speedup_vs_serial"],
        "unit": "x",
        "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": best["wall_s"] * 1e3,
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "fp32 transforms / fp64 state (fine), bf16 MFMA training, split-bf16 MFMA inference",
        "data": "synthetic",
        "config": {"workload": f"shoal basin {n}x{n}, 100 m square, depth 1.0 m, Gaussian shoal to "
                               f"{args.pr_shoal_depth} m (sigma 15 m), two reference depths, no influx, breaking off; "
                               f"sea Hs 0.04 Tp 2 s; {nslices} slices of {dtc} s; FNO 4 layers, 32 modes, width 32, "
                               f"GELU, hidden 128", "snapshots": nsnap, "training_pairs": int(steps / args.pr_epochs * 8),
                   "epochs": args.pr_epochs},
        "serial_fine_s": t_serial,
        "training": {"snapshot_s": t_snap, "train_s": t_train, "adam_steps": steps, "loss_first_epoch": hist[0],
                     "loss_last_epoch": hist[-1], "holdout_loss": val,
                     "gemm_tflops_over_steps": train_flops / t_train / 1e12},
        "roofline": {"bound": "mfma", "achieved": train_flops / t_train / 1e12, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": train_flops / t_train / 1e12 / 2500.0, "traffic": None,
                     "unit_of_work": "matrix-product flops of the training steps over the steps' wall time "
                                     "(all kernels and host calls included)"},
        "runs": runs,
    }
    print(json.dumps(out))


def bench_kin1d(args):
    """hw1d_kinematics (kinematic_modul, 1d/kinematic_modul.c) at one output
    time: 4096 nodes over a sloping bottom (14 -> 6 m), the reference's
    default 50 surface + 40 deep levels, a 2560-node window; each call takes
    host arrays and returns the six fields (PCIe and the call's own device
    allocation included).  CPU baseline: the numpy oracle on the same input."""
    import numpy as np
    from hawassi_c_amd.kinematics1d import Kinematics1D
    n, dx = 4096, 2.0
    rng = np.random.default_rng(7)
    x = np.arange(n) * dx
    L = n * dx
    k = 2 * np.pi * np.fft.fftfreq(n, d=dx)
    depth = 10.0 - 4.0 * np.tanh((x - 0.5 * L) / (0.1 * L))

    def smooth(amp):
        v = np.zeros(n)
        for m in rng.integers(2, 40, size=8):
            v += amp * rng.uniform(-1, 1) * np.cos(2 * np.pi * m * x / L + rng.uniform(0, 2 * np.pi))
        return v

    s = Kinematics1D(x, depth, k, 0.9, -0.5, 0.5, 50, 40, x[n // 8], x[n // 8 + 2559])
    eta, uh, deta, dph = smooth(0.05), np.fft.fft(smooth(0.3)), smooth(0.1), np.fft.fft(smooth(0.5))
    for _ in range(args.warmup):
        s.run(eta, uh, deta, dph, 1.0, 0.5)
    w0 = time.perf_counter()
    for _ in range(args.steps):
        s.run(eta, uh, deta, dph, 1.0, 0.5)
    wall = time.perf_counter() - w0
    out = {
        "metric": "1-D kinematics evaluations/s (kinematic_modul at one output time)",
        "value": args.steps / wall, "unit": "calls/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "kin1d: 4096 nodes, sloping bottom 14 -> 6 m, %d levels, %d-node window, host arrays "
                               "in and out" % (s.z.size, s.nxout)},
        "roofline": None,
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from oracle.kinematics1d_oracle import kinematics
        c0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            kinematics(s, eta, uh, deta, dph, 1.0, 0.5)
        cw = (time.perf_counter() - c0) / reps
        out["cpu_baseline"] = {"value": 1.0 / cw, "unit": "calls/s", "cores": 1, "kind": "port",
                               "sample": "3 calls of the numpy oracle on the same input"}
    print(json.dumps(out))


def bench_slope1d(args):
    """1-D batch: nreal realizations of the sloping-beach HS2 case (two
    reference depths, Up2IP, as the configuration names; breaking) with n drawn
    from {4096, 8192, 16384}, FP32 transforms / FP64 state, advanced together
    by the device driver through the reference's output times (every 0.05 s).
    A step is one output interval of the whole batch; value = realizations x
    simulated seconds per wall second.  Progress is printed every output
    interval; --round-limit bounds one interval's rounds (the reference's
    step control rejects steps in cascades in
    breaking realizations (the FP64 oracle needs up to ~1400 right-hand sides
    per 0.05 s for one of them, with steps down to 1e-6 s), and the batch
    advances at the pace of its slowest realization."""
    from hawassi_c_amd import HW_PREC_FP32
    from hawassi_c_amd import workloads1d as W1
    from hawassi_c_amd.model1d import Hawassi1D

    dt_out = 0.05
    spin = args.spinup if args.spinup != 150.0 else 2.0
    if args.steps == 20 and args.warmup == 10:
        args.steps, args.warmup = 10, 1
    t_end = spin + (args.warmup + args.steps + 4) * dt_out
    log(f"setup {args.nreal} realizations, {args.depths} reference depths")
    models = W1.slope_batch(args.nreal, seed=7, t_end=t_end, ndepth=args.depths)
    ns = [m.n for m in models]
    dev = Hawassi1D(models, HW_PREC_FP32, t0=0.0, h0=1e-4)
    dev.set_round_limit(args.round_limit)
    t = 0.0
    w0 = time.perf_counter()

    def advance(t_to):
        dev.integrate(t_to)
        return t_to

    while t < spin - 1e-12:
        t = advance(min(spin, round(t + dt_out, 10)))
        log(f"  spin-up t={t:.2f} wall {time.perf_counter() - w0:.1f} s")
    log(f"spin-up done t={t:.2f} stats={dev.stats()}")
    for _ in range(args.warmup):
        t = advance(round(t + dt_out, 10))
    st0 = dev.stats()
    t_start = t
    w0 = time.perf_counter()
    for i in range(args.steps):
        t = advance(round(t + dt_out, 10))
        st = dev.stats()
        log(f"  timed {i + 1}/{args.steps} t={t:.2f} wall {time.perf_counter() - w0:.1f} s "
            f"rhs {st['rhs_evals'] - st0['rhs_evals']} rejected {st['steps_rejected'] - st0['steps_rejected']}")
    wall = time.perf_counter() - w0
    st1 = dev.stats()
    sim = t - t_start
    nrhs = st1["rhs_evals"] - st0["rhs_evals"]
    log(f"timed {args.steps} output intervals: wall {wall:.3f} s, sim {sim:.3f} s, rhs {nrhs}")
    # algorithmic bytes per realization-RHS (transform field traffic model,
    # T_alg = 20 with two reference depths (23 with three), plus the
    # coefficient fields and the FP64 state in/out)
    t_alg = 20 if args.depths == 2 else 23

    def b_rhs(n):
        h = n // 2 + 1
        return t_alg * (4 * n + 8 * h) + 8 * 4 * n + 2 * (2 * 16 * h)
    batch_b = float(sum(b_rhs(n) for n in ns))
    mean_b = batch_b / len(ns)
    dt_launch = (st1["rhs_ms_total"] - st0["rhs_ms_total"]) / max(st1["rhs_timed"] - st0["rhs_timed"], 1)
    achieved = batch_b / (dt_launch * 1e-3) / 1e9
    job = nrhs * mean_b / wall / 1e9
    out = {
        "metric": "sim-sec/wall-sec (1-D batch, realizations x simulated seconds)",
        "value": args.nreal * sim / wall,
        "unit": "sim-s/wall-s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"slope1d batch of {args.nreal}: n in {{4096, 8192, 16384}} "
                               f"(counts {ns.count(4096)}/{ns.count(8192)}/{ns.count(16384)}), 400 m, depth 1.0 -> 0.2 m, "
                               + ("two reference depths (Up2IP)" if args.depths == 2 else
                                  "three reference depths (Up3IP, the reference's 1-D default)")
                               + ", JONSWAP Hs 0.10 Tp 2.0, breaking KBC 0.85",
                   "sim_window_s": [t_start, t]},
        "rhs_evals_per_s": nrhs / wall,
        "realization_rhs_bytes_mean": mean_b,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": RHS_TRAFFIC_BYTES_SLOPE1D if (args.nreal == 256 and args.depths == 2) else None,
                     "unit_of_work": "one right-hand side of the whole batch (the size classes' k1d_rhs_n "
                                     "launches of one stage side by side, every realization active), "
                                     "%.1f MB algorithmic, %.3f ms event-timed (the slowest class)"
                                     % (batch_b / 1e6, dt_launch)},
        "roofline_job": {"achieved": job, "frac": job / HBM_PEAK_GBS, "unit": "GB/s"},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from oracle import oracle as O
        r = ns.index(min(ns))
        m = models[r]
        y = dev.get_state(r)
        tr, hr = dev.time(r)
        orc = O.Oracle1D(m)
        c0 = orc.rhs_count()
        w0 = time.perf_counter()
        tt, hh = tr, hr
        for _ in range(4):
            rc, tt, hh, _ = orc.integrate(tt, tt + dt_out, hh, y)
        cw = time.perf_counter() - w0
        out["cpu_baseline"] = {"value": (tt - tr) / cw, "unit": "sim-s/wall-s per realization", "cores": 1,
                               "kind": "port",
                               "sample": f"4 output intervals ({tt - tr:.2f} s) of realization {r} (n={m.n}) from "
                                         f"the device state at t={tr:.2f}; FP64 oracle, 1 thread, {cpu_model()}",
                               "rhs_evals_per_s": (orc.rhs_count() - c0) / cw}
    dev.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
