"""Autotuner for the launch geometry of the fused 2-D kernels.

The fused half-spectrum kernels (csrc/rhs2d_kernels.h and friends) take their
block geometry at compile time, per (precision, grid) unit:

* HW_X_RUN_BYTES   -- rows per x-pass block (the transposed spectral runs
                      X[kx][y0 .. y0+RR) are this many bytes long);
* HW_X_PER_THREAD  -- elements per thread in the x-pass FFT, which also sets
                      its radix schedule (radix 8 passes for 8, radix 16 for 16);
* HW_Y_PER_THREAD  -- the same for the y-pass (and with it the columns a
                      y-pass block owns);
* HW_Y_WPE32       -- waves per SIMD the FP32 y-pass kernels are compiled for
                      (their register budget);
* HW_X_LDS_BUDGET,
  HW_X_WPE         -- the LDS and register budgets of an x-pass block: with
                      smaller ones several x blocks share a CU, and one
                      block's loads overlap another's LDS transform phases.

These macros are read only by XGeo<T, NX> / YGeo<T, NY> and the kernels
instantiated with them, so a unit whose sides occur in no other unit (the
1024^2 and 2048^2 squares) can take its own values without changing any
other unit's code.

Three steps, the middle one on the GPU:

  python -m hawassi_c_amd.autotune build  [--unit p1_2048x2048]
      compiles every candidate of the unit into its own library under
      hawassi_c_amd/_tune/<unit>/<candidate>/libhawassi_hip.so (the 2-D
      entry points and that unit only; a candidate whose geometry does not
      fit is recorded as such);
  python -m hawassi_c_amd.autotune run    [--unit ...] [--out FILE]
      times each candidate library in a child process: the model the unit
      serves (the flat open sea at 1024^2, the two-depth harbour otherwise), a
      random state, the event-timed average right-hand side over a few
      integrated steps; writes the table as JSON;
  python -m hawassi_c_amd.autotune apply  FILE [FILE ...]
      writes the fastest candidate of each unit to hawassi_c_amd/tuned.json,
      which build.py reads (its -D flags go to that unit only).

A second "build --center FILE" sweeps around the best of an earlier run
(one pass's geometry changed at a time), so two rounds combine the best
x-pass and the best y-pass geometry.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import itertools
import json
import os
import pickle
import subprocess
import sys
import tempfile
import time
from pathlib import Path

from . import build as B

TUNED = B.PKG / "tuned.json"
TUNE_DIR = B.PKG / "_tune"           # candidate libraries (travel to the GPU; git-ignored)

# the defaults of hw_fft.h, the first candidate of every unit
DEFAULT = {"HW_X_RUN_BYTES": 64, "HW_X_PER_THREAD": 8, "HW_Y_PER_THREAD": 8, "HW_Y_WPE32": 3}


def candidates(unit: str, center: dict | None = None, skip: list | None = None) -> list[dict]:
    """The geometry candidates of a unit around `center` (the defaults
    first): runs of 32/64/128 bytes x 8 or 16 elements per x thread, 8 or 16
    per y thread, 2/3/4 waves per SIMD in the y-pass, changing either the
    x-pass geometry or the y-pass geometry of the center (the two passes are
    separate kernels, so a second sweep around the best of the first
    combines the best of each); geometries a block cannot hold are dropped
    by the compiler's static checks and recorded as such."""
    c0 = dict(center or DEFAULT)
    out = [c0]
    for run, ex, ey, wpe in itertools.product((32, 64, 128), (8, 16), (8, 16), (2, 3, 4)):
        c = {"HW_X_RUN_BYTES": run, "HW_X_PER_THREAD": ex, "HW_Y_PER_THREAD": ey, "HW_Y_WPE32": wpe}
        if c == c0:
            continue
        dx = (run, ex) != (c0["HW_X_RUN_BYTES"], c0["HW_X_PER_THREAD"])
        dy = (ey, wpe) != (c0["HW_Y_PER_THREAD"], c0["HW_Y_WPE32"])
        if dx and dy:
            continue
        out.append(c)
    # several x blocks per CU: a smaller LDS budget (one exchange buffer or
    # twiddles from global memory) and a register budget for that many waves
    for ex, lds, wpe in ((8, 81920, 4), (8, 54613, 6), (16, 81920, 4)):
        c = {**c0, "HW_X_RUN_BYTES": 32, "HW_X_PER_THREAD": ex, "HW_X_LDS_BUDGET": lds, "HW_X_WPE": wpe}
        if c not in out:
            out.append(c)
    return [c for c in out if not skip or c not in skip]


def cand_name(c: dict) -> str:
    s = "x%d_e%d_y%d_w%d" % (c["HW_X_RUN_BYTES"], c["HW_X_PER_THREAD"], c["HW_Y_PER_THREAD"], c["HW_Y_WPE32"])
    if "HW_X_LDS_BUDGET" in c:
        s += "_l%d_o%d" % (c["HW_X_LDS_BUDGET"] // 1024, c["HW_X_WPE"])
    return s


def unit_source(unit: str) -> Path:
    return B.BUILD / "gen" / f"inst_{unit}.hip"


def unit_size(unit: str) -> tuple[int, int, int]:
    p, s = unit.split("_")
    nx, ny = s.split("x")
    return int(p[1:]), int(nx), int(ny)


def _build_one(unit: str, c: dict) -> dict:
    d = TUNE_DIR / unit / cand_name(c)
    d.mkdir(parents=True, exist_ok=True)
    obj = d / "unit.o"
    lib = d / "libhawassi_hip.so"
    flags = [f"-D{k}={v}" for k, v in c.items()]
    cmd = [B.HIPCC, *B.FLAGS, *flags, "-I", str(B.CSRC), "-I", str(B.PKG.parent / "include"), "-c",
           str(unit_source(unit)), "-o", str(obj)]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        last = [ln for ln in r.stderr.splitlines() if "error" in ln]
        return {"candidate": cand_name(c), "flags": c, "ok": False, "why": (last[0] if last else r.stderr[-300:])[:300]}
    # the 2-D entry points and this one unit: enough for the harbour model
    common = [B.BUILD / f"{s}.o" for s in ("capi", "full2d")]
    cmd = [B.HIPCC, f"--offload-arch={B.ARCH}", "-shared", "-fPIC", "-o", str(lib), *map(str, common), str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return {"candidate": cand_name(c), "flags": c, "ok": False, "why": "link: " + r.stderr[-300:]}
    return {"candidate": cand_name(c), "flags": c, "ok": True, "lib": str(lib.relative_to(B.PKG.parent)),
            "compile_s": round(time.time() - t0, 1)}


def cmd_build(units: list[str], jobs: int, center_file: str | None = None) -> None:
    B.build(verbose=False)             # the common objects and the generated unit sources
    f = TUNE_DIR / "candidates.json"
    rows = [r for r in json.loads(f.read_text()) if r["unit"] not in units] if f.exists() else []
    prev = json.loads(Path(center_file).read_text()) if center_file else []
    for unit in units:
        timed = [r for r in prev if r["unit"] == unit and r.get("rhs_us") and r.get("same_state")]
        center = min(timed, key=lambda r: r["rhs_us"])["flags"] if timed else None
        # a second sweep: around the best of the first, without what it timed
        cs = candidates(unit, center, [r["flags"] for r in timed] if center else None)
        with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
            for res in ex.map(lambda c: _build_one(unit, c), cs):
                res["unit"] = unit
                rows.append(res)
                print(("built  " if res["ok"] else "skip   ") + unit + " " + res["candidate"] +
                      ("" if res["ok"] else "  (" + res["why"] + ")"), flush=True)
    (TUNE_DIR / "candidates.json").write_text(json.dumps(rows, indent=1))


# ------------------------------------------------------------ GPU side ----
def _child(lib_path: str, model_file: str, steps: float, partial: bool = True) -> None:
    """Times one candidate library (runs in its own process)."""
    import numpy as np

    from . import _lib
    from .model2d import Hawassi2D
    _lib._lib = _lib.load(Path(lib_path), partial=partial)   # every later load() returns the candidate
    with open(model_file, "rb") as f:
        m, y0, prec = pickle.load(f)
    d = Hawassi2D(m, prec)
    d.set_state(y0)
    t, h = d.integrate(0.0, 0.3, 0.02)        # warm-up: graphs of nothing, caches, clocks
    d.reset_stats()
    t, h = d.integrate(t, t + steps, h)
    st = d.stats()
    y = d.get_state()
    d.close()
    out = {"rhs_us": 1e3 * st["rhs_ms_total"] / max(1, st["rhs_timed"]), "rhs_evals": st["rhs_evals"],
           "rhs_timed": st["rhs_timed"], "state_norm": float(np.linalg.norm(y))}
    print("RESULT " + json.dumps(out), flush=True)


def cmd_run(units: list[str], out_file: str, steps: float) -> None:
    import numpy as np

    from . import workloads as W
    rows = json.loads((TUNE_DIR / "candidates.json").read_text())
    results = []
    for unit in units:
        prec, nx, ny = unit_size(unit)
        # the model the unit serves: the flat open sea at 1024^2 (2-D batch),
        # the two-depth harbour otherwise
        kind = W.short_crested if unit == "p1_1024x1024" else W.harbour
        m = kind(nx, t_end=30.0) if nx == ny else kind(nx, ny=ny, t_end=30.0)
        y0 = W.random_state(m, 7, amp=0.5, kfrac=0.3)
        mf = tempfile.NamedTemporaryFile(suffix=".pkl", delete=False)
        pickle.dump((m, y0, prec), mf)
        mf.close()
        ref = None
        # the library as built (its tuned.json flags) first, for reference
        main = {"unit": unit, "candidate": "built", "flags": None, "ok": True,
                "lib": str(B.LIB.relative_to(B.PKG.parent))}
        for r in [main] + rows:
            if r["unit"] != unit or not r["ok"]:
                continue
            print(f"tune {unit} {r['candidate']}", flush=True)
            p = subprocess.run([sys.executable, "-m", "hawassi_c_amd.autotune", "child", str(B.PKG.parent / r["lib"]),
                                mf.name, str(steps), "full" if r["candidate"] == "built" else "partial"],
                               capture_output=True, text=True, timeout=300, cwd=str(B.PKG.parent))
            line = [ln for ln in p.stdout.splitlines() if ln.startswith("RESULT ")]
            if p.returncode != 0 or not line:
                results.append({**r, "rhs_us": None, "why": (p.stderr or p.stdout)[-300:]})
                print("  failed", flush=True)
                continue
            res = json.loads(line[0][7:])
            if ref is None:
                ref = res["state_norm"]
            # every geometry computes the same right-hand side
            res["same_state"] = abs(res["state_norm"] - ref) <= 1e-5 * ref
            results.append({**r, **res})
            print(f"  {res['rhs_us']:.1f} us per RHS ({res['rhs_timed']} timed)", flush=True)
        os.unlink(mf.name)
    Path(out_file).parent.mkdir(parents=True, exist_ok=True)
    Path(out_file).write_text(json.dumps(results, indent=1))


def cmd_apply(result_files: list[str]) -> None:
    res = [r for f in result_files for r in json.loads(Path(f).read_text())]
    tuned = json.loads(TUNED.read_text()) if TUNED.exists() else {}
    for unit in sorted({r["unit"] for r in res}):
        ok = [r for r in res if r["unit"] == unit and r.get("rhs_us") and r.get("same_state") and r.get("flags")]
        if not ok:
            continue
        best = min(ok, key=lambda r: r["rhs_us"])
        default = next((r for r in ok if r["flags"] == DEFAULT), None)
        tuned[unit] = {"flags": best["flags"], "rhs_us": round(best["rhs_us"], 1),
                       "default_rhs_us": round(default["rhs_us"], 1) if default else None,
                       "candidates_timed": len(ok)}
        print(f"{unit}: {best['candidate']} {best['rhs_us']:.1f} us"
              + (f" (default {default['rhs_us']:.1f} us)" if default else ""))
    TUNED.write_text(json.dumps(tuned, indent=1, sort_keys=True) + "\n")


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(prog="python -m hawassi_c_amd.autotune")
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build")
    b.add_argument("--unit", action="append", default=None)
    b.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 4))
    b.add_argument("--center", default=None, help="results of an earlier run: sweep around its best candidate")
    r = sub.add_parser("run")
    r.add_argument("--unit", action="append", default=None)
    r.add_argument("--out", default="gpurun_out/tune.json")
    r.add_argument("--steps", type=float, default=1.0, help="simulated seconds timed per candidate")
    a = sub.add_parser("apply")
    a.add_argument("file", nargs="+")
    c = sub.add_parser("child")
    c.add_argument("lib")
    c.add_argument("model")
    c.add_argument("steps", type=float)
    c.add_argument("mode", choices=["full", "partial"])
    args = ap.parse_args(argv)
    units = getattr(args, "unit", None) or ["p1_2048x2048"]
    if args.cmd == "build":
        cmd_build(units, args.jobs, args.center)
    elif args.cmd == "run":
        cmd_run(units, args.out, args.steps)
    elif args.cmd == "apply":
        cmd_apply(args.file)
    else:
        _child(args.lib, args.model, args.steps, args.mode == "partial")


if __name__ == "__main__":
    main()
