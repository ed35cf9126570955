"""Build libgraphhip.so in-tree: hipcc for the gfx950 kernels, g++ for the
host-side graph construction, one shared library next to this file.  Then
g++ builds tests/cpp/dropin_check, the C++ program that drives the drop-in
headers (include/graphkokkos/*.h) against that library.

    python -m graphkokkos_amd.build            # build (always recompiles)
    python -m graphkokkos_amd.build --check    # report whether it is stale
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
OUT = PKG / "libgraphhip.so"
BUILD_DIR = PKG / "_build"
DROPIN_SRC = ROOT / "tests" / "cpp" / "dropin_check.cpp"
DROPIN_BIN = ROOT / "tests" / "cpp" / "_build" / "dropin_check"

HIP_SOURCES = ["gk_runtime.hip", "spgemm.hip", "ktruss.hip", "bc.hip", "selector.hip", "batch.hip", "graph_device.hip", "cmask.hip",
               "tc_vertex.hip", "bc_noalloc.hip"]
HOST_SOURCES = ["host_graph.cpp"]
ARCH = os.environ.get("GK_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; a ROCm installation is required to build libgraphhip.so")


def _hip_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        # packed FP32 instructions stay disabled for every kernel
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{INCLUDE}",
        f"-I{CSRC}",
        "-Wno-unused-result",
    ]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def sources() -> list[Path]:
    return [CSRC / s for s in HIP_SOURCES + HOST_SOURCES] + sorted(CSRC.glob("*.h")) + sorted(INCLUDE.glob("**/*.h"))


def is_stale() -> bool:
    if not OUT.exists():
        return True
    t = OUT.stat().st_mtime
    return any(p.stat().st_mtime > t for p in sources())


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    hipcc = _hipcc()
    BUILD_DIR.mkdir(exist_ok=True)
    objs: list[Path] = []
    cmds: list[list[str]] = []
    for s in HIP_SOURCES:
        o = BUILD_DIR / (Path(s).stem + ".o")
        cmds.append([hipcc, *_hip_flags(), "-c", str(CSRC / s), "-o", str(o)])
        objs.append(o)
    for s in HOST_SOURCES:
        o = BUILD_DIR / (Path(s).stem + ".o")
        cmds.append(["g++", "-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wno-sign-compare",
                     f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(CSRC / s), "-o", str(o)])
        objs.append(o)
    jobs = jobs or min(len(cmds), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for cmd, fut in [(c, ex.submit(_run, c)) for c in cmds]:
            if verbose:
                print(" ".join(cmd), flush=True)
            fut.result()
    tmp = OUT.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-pthread"])
    tmp.replace(OUT)
    build_dropin_check()
    return OUT


def build_dropin_check() -> Path:
    """The drop-in headers compiled as a client of libgraphhip.so would."""
    DROPIN_BIN.parent.mkdir(parents=True, exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Wno-sign-compare", f"-I{INCLUDE}", str(DROPIN_SRC),
          "-o", str(DROPIN_BIN), f"-L{PKG}", "-lgraphhip", "-Wl,-rpath,$ORIGIN/../../../graphkokkos_amd"])
    return DROPIN_BIN


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--check", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    if a.check:
        print("stale" if is_stale() else "fresh")
        return 0
    print(build(verbose=a.verbose))
    return 0


if __name__ == "__main__":
    sys.exit(main())
