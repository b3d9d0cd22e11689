"""In-tree build of libvfs.so (gfx950) and of the test-side native pieces.

    python -m vfs_amd.build          # everything
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")

ARCH = os.environ.get("VFS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# Packed-FP32 instruction forms are disabled for all device code.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# Translation units of libvfs.so: the launchers and the host solver, plus the op kernels
# instantiated once per storage type (0 float, 1 __half, 2 __hip_bfloat16) and kernel
# family (1 = tick kernel; 2, 4, 8 = K-iteration kernel; 0 = whole-warp persistent kernel)
# from tgv_inst.hip, compiled in parallel into objects under vfs_amd/_build/.
LIB_SOURCES = ["kernels.hip", "solver.hip"]
INST_UNITS = [(st, k) for st in (0, 1, 2) for k in (1, 2, 4, 8, 0)]
BUILD_DIR = os.path.join(PKG, "_build")


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def _jobs():
    return max(1, min(int(os.environ.get("MAX_JOBS", os.cpu_count() or 1)), 16))


def build_lib(verbose=False) -> str:
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(BUILD_DIR, exist_ok=True)
    compiles = []
    for src in LIB_SOURCES:
        obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
        compiles.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    for st, k in INST_UNITS:
        obj = os.path.join(BUILD_DIR, f"tgv_inst_s{st}_k{k}.o")
        compiles.append([HIPCC] + HIP_FLAGS + [f"-DVFS_INST_STORAGE={st}", f"-DVFS_INST_K={k}", "-c",
                                              os.path.join(CSRC, "tgv_inst.hip"), "-o", obj])
    # the slowest units (multi-iteration kernels) first
    compiles.sort(key=lambda c: 0 if any(f"-DVFS_INST_K={k}" in c for k in (0, 2, 4, 8)) else 1)
    with ThreadPoolExecutor(_jobs()) as ex:
        list(ex.map(lambda c: _run(c, verbose), compiles))
    out = os.path.join(PKG, "libvfs.so")
    tmp = out + ".tmp"
    objs = [c[c.index("-o") + 1] for c in compiles]
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, out)
    return out


def build_dllhook(verbose=False) -> str:
    """libvfs_dllhook.so: the reference's DLL entry points (Create / Init / SolveStereo)
    over libvfs (host C++ only)."""
    out = os.path.join(PKG, "libvfs_dllhook.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "dllhook.cpp"), "-o", tmp, "-L", PKG, "-lvfs", "-lz", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, out)
    return out


def build_dropin(verbose=False) -> str | None:
    """The C++ consumer that replays test_vfs/main.cpp through include/stereotgv/stereotgv.h."""
    src = os.path.join(ROOT, "tests", "dropin", "test_vfs_dropin.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "dropin", "test_vfs_dropin")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-o", out,
          "-L", PKG, "-lvfs", "-lz", "-Wl,-rpath,$ORIGIN/../../vfs_amd"], verbose)
    return out


def build_order_check(verbose=False) -> str | None:
    """Host program that checks the persistent op's work-item order (vfs_engine.h)."""
    src = os.path.join(ROOT, "tests", "persist_order", "check_persist_order.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "persist_order", "check_persist_order")
    _run([HIPCC, "-O1", "-std=c++17", "-Wall", src, "-o", out], verbose)
    return out


def build_all(verbose=False):
    build_lib(verbose)
    build_dllhook(verbose)
    build_dropin(verbose)
    build_order_check(verbose)


if __name__ == "__main__":
    build_all(verbose="-q" not in sys.argv)
