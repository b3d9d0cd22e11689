"""libvfs_dllhook.so: the reference's DLL entry points (dllhook/vfs.h:38-45,
dllhook/vfs.cpp:3-57) over libvfs.  Init reads the fisheye mask (8-bit PNG) and the
.flo fields from files, SolveStereo equalizes both 8-bit frames, solves, and writes the
8-bit depth map; its disparity must equal the Python mirror's on the same inputs."""
import ctypes
import os
import re

import numpy as np
import pytest

from vfs_amd import imageio as io
from vfs_amd import synth as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "vfs_amd", "libvfs_dllhook.so")
D = os.path.join(ROOT, "tests", "golden", "test_vfs")


def declared():
    hdr = open(os.path.join(ROOT, "include", "vfs_dllhook.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z]+)\s*\(", hdr)) - {"defined"})


def load():
    L = ctypes.CDLL(LIB)
    L.Create.restype = ctypes.c_void_p
    L.Init.argtypes = ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 6
                       + [ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float]
                       + [ctypes.c_char_p] * 3)
    L.SolveStereo.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_ubyte)] * 3 + [ctypes.c_int, ctypes.c_int,
                                                                                          ctypes.c_float]
    L.VfsDllLastError.restype = ctypes.c_char_p
    L.VfsDllLastError.argtypes = [ctypes.c_void_p]
    L.VfsDllDestroy.argtypes = [ctypes.c_void_p]
    return L


def test_dllhook_exports_the_reference_entry_points():
    assert os.path.exists(LIB), "libvfs_dllhook.so not built"
    names = declared()
    assert {"Create", "Init", "SolveStereo"} <= set(names)
    L = ctypes.CDLL(LIB)
    for n in names:
        assert hasattr(L, n), n


def _u8p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))


@pytest.mark.gpu
def test_dllhook_solve_matches_mirror():
    from vfs_amd import StereoTgv
    P = S.SOLVER_PARAMS
    L = load()
    W, H, LV, FS, WI, IT = 424, 400, 5, 2.0, 20, 10
    focal, baseline, max_depth = 143.259, 0.0639505, 6.0
    f1 = np.ascontiguousarray(io.resize_half(io.read_png_gray(os.path.join(D, "wang1.png"))))
    f2 = np.ascontiguousarray(io.resize_half(io.read_png_gray(os.path.join(D, "wang2.png"))))
    out = np.zeros((H, W), np.uint8)
    h = L.Create()
    assert h
    # SolveStereo before Init reports an error instead of exiting
    L.SolveStereo(h, _u8p(f1), _u8p(f2), _u8p(out), W, H, max_depth)
    assert L.VfsDllLastError(h) != b""
    print("dllhook: Init + SolveStereo 424x400 on the reference's wang fixtures", flush=True)
    L.Init(h, W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], LV, FS, WI, IT,
           focal, baseline, os.path.join(D, "maskHalf.png").encode(),
           os.path.join(D, "translationVectorWangHalf.flo").encode(),
           os.path.join(D, "calibrationVectorWangHalf.flo").encode())
    assert L.VfsDllLastError(h) == b"", L.VfsDllLastError(h)
    L.SolveStereo(h, _u8p(f1), _u8p(f2), _u8p(out), W, H, max_depth)
    assert L.VfsDllLastError(h) == b"", L.VfsDllLastError(h)

    mask = io.read_png_gray(os.path.join(D, "maskHalf.png")).astype(np.float32) * np.float32(1.0 / 255.0)
    g = StereoTgv()
    g.initialize(W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], LV, FS, WI, IT)
    g.copyMaskToDevice(mask)
    g.loadVectorFields(io.read_flo(os.path.join(D, "translationVectorWangHalf.flo")),
                       io.read_flo(os.path.join(D, "calibrationVectorWangHalf.flo")))
    g.copyImagesToDevice(io.equalize_hist(f1), io.equalize_hist(f2))
    g.solveStereoForwardMasked()
    u = g.copy1DDisparityToHost()
    fb = np.float32(focal) * np.float32(baseline)
    with np.errstate(divide="ignore"):
        z = np.minimum(np.where(u > 0, fb / np.maximum(u, np.float32(1e-30)), 0), np.float32(max_depth))
    ref = np.where(u > 0, np.rint(np.float32(255) * z.astype(np.float32) / np.float32(max_depth)), 0).astype(np.uint8)
    diff = np.abs(out.astype(int) - ref.astype(int))
    assert diff.max() <= 1 and (diff == 0).mean() > 0.999, (diff.max(), (diff == 0).mean())
    # a plausible depth map: most of the disc has finite depth below the cap
    m = mask > 0
    assert ((out[m] > 0) & (out[m] < 255)).mean() > 0.5
    # Init with a missing file reports it
    L.Init(h, W, H, P["beta"], P["gamma"], P["alpha0"], P["alpha1"], P["timestep_lambda"], P["lam"], LV, FS, WI, IT,
           focal, baseline, b"/nonexistent/mask.png", b"/nonexistent/tv.flo", b"/nonexistent/cv.flo")
    assert b"mask" in L.VfsDllLastError(h)
    L.VfsDllDestroy(h)
