"""Build libhawassi_hip.so for gfx950 with hipcc (in-tree).

One translation unit per (precision, nx, ny) kernel set is generated from the
HW_INSTANTIATE macro and compiled in parallel; the C ABI unit (capi.hip) is
linked with them into hawassi_c_amd/libhawassi_hip.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import json
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libhawassi_hip.so"

ARCH = os.environ.get("HW_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]

# (precision, nx, ny): precision 0 = FP64, 1 = FP32 transforms/operators
SIZES = (
    [(0, n, n) for n in (32, 64, 128, 256, 512, 1024, 2048)]
    + [(1, n, n) for n in (32, 64, 128, 256, 512, 1024, 2048)]
    + [(p, a, b) for p in (0, 1) for (a, b) in ((64, 32), (32, 64), (128, 64), (256, 64), (512, 256))]
)


# 1-D batched right-hand side: one unit per (precision, size class n), the
# step-control kernel per (precision, workgroup size); n = 2000 is the grid
# of the reference case 1dBBr001HarmBar (1000 threads)
N1D = (128, 256, 512, 1024, 2000, 2048, 4096, 8192, 16384)
NT1D = (64, 128, 256, 512, 1000, 1024)


def _units():
    gen = BUILD / "gen"
    gen.mkdir(parents=True, exist_ok=True)
    units = [CSRC / "capi.hip", CSRC / "capi1d.hip", CSRC / "fno.hip", CSRC / "full2d.hip", CSRC / "kin1d.hip"]
    for prec in (0, 1):
        t = "double" if prec == 0 else "float"
        for n in N1D:
            if n * (16 if prec == 0 else 8) > 131072:
                continue      # one complex field of the realization in LDS
            f = gen / f"inst1d_p{prec}_n{n}.hip"
            k = f"k1d_rhs_n<{t}, nt_for_n({n}), {n}"
            text = (f'// generated by hawassi_c_amd/build.py\n#include "{CSRC}/solver1d.h"\n'
                    f"namespace hw {{\n"
                    f"hipError_t k1d_launch_{prec}_{n}(dim3 g, size_t lds, hipStream_t s, const Batch1DArgs<{t}>& a, "
                    f"int st, int list0, int mid) {{\n"
                    f"    if (mid) hipLaunchKernelGGL(({k}, true>), g, dim3(nt_for_n({n})), lds, s, a, st, list0);\n"
                    f"    else hipLaunchKernelGGL(({k}, false>), g, dim3(nt_for_n({n})), lds, s, a, st, list0);\n"
                    f"    return hipGetLastError();\n}}\n"
                    f"hipError_t k1d_setlds_{prec}_{n}(size_t b) {{\n"
                    f"    hipError_t e = hipFuncSetAttribute((const void*){k}, true>, "
                    f"hipFuncAttributeMaxDynamicSharedMemorySize, (int)b);\n"
                    f"    if (e != hipSuccess) return e;\n"
                    f"    return hipFuncSetAttribute((const void*){k}, false>, "
                    f"hipFuncAttributeMaxDynamicSharedMemorySize, (int)b);\n}}\n"
                    f"}}  // namespace hw\n")
            if not f.exists() or f.read_text() != text:
                f.write_text(text)
            units.append(f)
        f = gen / f"inst1d_p{prec}_ctl.hip"
        text = f'// generated by hawassi_c_amd/build.py\n#include "{CSRC}/solver1d.h"\nnamespace hw {{\n'
        for nt in NT1D:
            text += (f"hipError_t k1d_control_{prec}_{nt}(dim3 g, hipStream_t s, const Batch1DArgs<{t}>& a, int list0) {{\n"
                     f"    hipLaunchKernelGGL((k1d_control<{t}, {nt}>), g, dim3({nt}), 0, s, a, list0);\n"
                     f"    return hipGetLastError();\n}}\n")
        text += "}  // namespace hw\n"
        if not f.exists() or f.read_text() != text:
            f.write_text(text)
        units.append(f)
    for prec, nx, ny in SIZES:
        t = "double" if prec == 0 else "float"
        f = gen / f"inst_p{prec}_{nx}x{ny}.hip"
        text = (f'// generated by hawassi_c_amd/build.py\n#include "{CSRC}/solver2d.h"\n'
                f"HW_INSTANTIATE({t}, {prec}, {nx}, {ny})\n")
        if not f.exists() or f.read_text() != text:
            f.write_text(text)
        units.append(f)
    return units


def _includes(src: Path, seen: set) -> None:
    """The project headers src includes, transitively (quoted includes)."""
    for line in src.read_text().splitlines():
        line = line.strip()
        if not line.startswith("#include \""):
            continue
        name = line.split('"')[1]
        for base in (src.parent, CSRC, PKG.parent / "include"):
            f = (base / name).resolve()
            if f.exists():
                if f not in seen:
                    seen.add(f)
                    _includes(f, seen)
                break


def _tuned_flags(src: Path) -> list[str]:
    """-D flags the autotuner chose for this unit (hawassi_c_amd/tuned.json;
    see autotune.py), none for units it has not tuned."""
    tuned = PKG / "tuned.json"
    if not tuned.exists() or not src.stem.startswith("inst_p"):
        return []
    t = json.loads(tuned.read_text()).get(src.stem[len("inst_"):])
    return [f"-D{k}={v}" for k, v in sorted(t["flags"].items())] if t else []


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    seen: set = set()
    _includes(src, seen)
    deps = [src] + sorted(seen)
    tflags = _tuned_flags(src)
    if tflags:
        deps.append(PKG / "tuned.json")
    if obj.exists() and all(obj.stat().st_mtime >= d.stat().st_mtime for d in deps):
        return obj
    cmd = [HIPCC, *FLAGS, *tflags, "-I", str(CSRC), "-I", str(PKG.parent / "include"), "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr[-4000:]}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(parents=True, exist_ok=True)
    units = _units()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, units))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} ({len(objs)} units)")
    return LIB


if __name__ == "__main__":
    build(int(sys.argv[1]) if len(sys.argv) > 1 else None)
