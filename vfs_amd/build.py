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

LIB_SOURCES = ["kernels.hip", "solver.hip"]


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build_lib(verbose=False) -> str:
    out = os.path.join(PKG, "libvfs.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES]
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-o", tmp] + srcs, verbose)
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


def build_all(verbose=False):
    build_lib(verbose)
    build_dllhook(verbose)
    build_dropin(verbose)


if __name__ == "__main__":
    build_all(verbose="-q" not in sys.argv)
