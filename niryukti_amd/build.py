"""Builds the engine's native library in-tree for gfx950.

    python -m niryukti_amd.build [-v]

Host C++20 sources are compiled with g++, the HIP kernels with hipcc for
gfx950, and everything is linked by hipcc into niryukti_amd/lib/libniryukti.so
(plus the `niryukti` command-line tool in niryukti_amd/bin/).
"""
from __future__ import annotations

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
LIB_DIR = PKG / "lib"
BIN_DIR = PKG / "bin"
OBJ_DIR = PKG / "lib" / "obj"
LIBRARY = LIB_DIR / "libniryukti.so"
ARCH = os.environ.get("NIRYUKTI_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = sorted((CSRC / "engine").glob("*.cpp"))
HIP_SOURCES = sorted((CSRC / "hip").glob("*.hip"))
CLI_SOURCES = sorted((CSRC / "cli").glob("*.cpp"))

HOST_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-Wall", "-Wextra", f"-I{INCLUDE}", f"-I{CSRC}", f"-I{CSRC / 'engine'}"]
HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}", f"-I{INCLUDE}", f"-I{CSRC}",
    # Packed-FP32 instruction selection is disabled for all device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the engine needs ROCm's hipcc to build its gfx950 kernels")


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    proc = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(map(str, cmd))}\n{proc.stdout}\n{proc.stderr}")


def build(verbose: bool = False) -> Path:
    """Compiles every source (always, also when a library exists) and links it."""
    hipcc = _hipcc()
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    BIN_DIR.mkdir(parents=True, exist_ok=True)
    jobs = []
    objects = []
    for src in HOST_SOURCES:
        obj = OBJ_DIR / (src.stem + ".o")
        objects.append(obj)
        jobs.append(["g++", *HOST_FLAGS, "-c", src, "-o", obj])
    for src in HIP_SOURCES:
        obj = OBJ_DIR / (src.stem + ".hip.o")
        objects.append(obj)
        jobs.append([hipcc, *HIP_FLAGS, "-c", src, "-o", obj])
    workers = min(8, len(jobs), os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=workers) as pool:
        for fut in [pool.submit(_run, j, verbose) for j in jobs]:
            fut.result()
    tmp = LIBRARY.with_suffix(".so.tmp")
    _run([hipcc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objects, "-lpthread"], verbose)
    os.replace(tmp, LIBRARY)
    if CLI_SOURCES:
        cli = BIN_DIR / "niryukti"
        _run(["g++", *HOST_FLAGS, *CLI_SOURCES, "-o", cli, f"-L{LIB_DIR}", "-lniryukti",
              "-Wl,-rpath,$ORIGIN/../lib", "-lpthread"], verbose)
    return LIBRARY


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
