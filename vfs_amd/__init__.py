"""Python host mirror of the StereoTgv class surface over libvfs's C ABI.

Method names, argument meaning and order follow the reference class
(menandro/vfs stereotgv/stereotgv.h:107-124) so that scripts read like
test_vfs/main.cpp.  All compute runs in libvfs.so (HIP kernels for gfx950); there is
no CPU fallback: if the library is missing or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = ["StereoTgv", "Batch", "Server", "VfsError", "lib", "lib_path", "EXPORTS", "STORAGE"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_u8_p = ctypes.POINTER(ctypes.c_uint8)
_c_int_p = ctypes.POINTER(ctypes.c_int)

# (name, restype, argtypes) for every entry point declared in include/vfs.h
EXPORTS = [
    ("vfs_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    ("vfs_create_ex", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("vfs_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("vfs_initialize", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 6
     + [ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_int]),
    ("vfs_set_limit_range", ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
    ("vfs_get_limit_range", ctypes.c_int, [ctypes.c_void_p, _c_float_p]),
    ("vfs_copy_mask", ctypes.c_int, [ctypes.c_void_p, _c_float_p, ctypes.c_int]),
    ("vfs_load_vector_fields", ctypes.c_int, [ctypes.c_void_p, _c_float_p, ctypes.c_int, _c_float_p, ctypes.c_int]),
    ("vfs_copy_images_u8", ctypes.c_int, [ctypes.c_void_p, _c_u8_p, _c_u8_p, ctypes.c_int]),
    ("vfs_copy_images_f32", ctypes.c_int, [ctypes.c_void_p, _c_float_p, _c_float_p, ctypes.c_int]),
    ("vfs_copy_images_u8c3", ctypes.c_int, [ctypes.c_void_p, _c_u8_p, _c_u8_p, ctypes.c_int]),
    ("vfs_solve_masked", ctypes.c_int, [ctypes.c_void_p]),
    ("vfs_solve", ctypes.c_int, [ctypes.c_void_p]),
    ("vfs_solve_warm", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 4),
    ("vfs_copy_disparity_1d", ctypes.c_int, [ctypes.c_void_p, _c_float_p, ctypes.c_int]),
    ("vfs_copy_flow", ctypes.c_int, [ctypes.c_void_p, _c_float_p, ctypes.c_int]),
    ("vfs_copy_stereo", ctypes.c_int, [ctypes.c_void_p, _c_float_p, ctypes.c_int, _c_float_p, ctypes.c_int]
     + [ctypes.c_float] * 11),
    ("vfs_copy_warped_image", ctypes.c_int, [ctypes.c_void_p, _c_float_p, ctypes.c_int]),
    ("vfs_copy_warped_image_noalloc", ctypes.c_int, [ctypes.c_void_p, _c_float_p, ctypes.c_int]),
    ("vfs_copy_disparity_vis", ctypes.c_int, [ctypes.c_void_p, _c_float_p, ctypes.c_int, ctypes.c_float]),
    ("vfs_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
    ("vfs_get_level_dims", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p]),
    ("vfs_read_plane", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _c_float_p]),
    ("vfs_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("vfs_get_energy", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    ("vfs_set_option", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]),
    ("vfs_time_solves", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_float_p]),
    ("vfs_time_inner_iterations", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_float_p, _c_int_p]),
    ("vfs_time_copy", ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, _c_float_p]),
    ("vfs_get_option", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _c_int_p]),
    ("vfs_get_level_tuning", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p,
                                            _c_int_p]),
    ("vfs_set_level_tuning", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 5),
    ("vfs_batch_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    ("vfs_batch_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("vfs_batch_solve", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("vfs_batch_solve_warm", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 4),
    ("vfs_batch_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("vfs_batch_get_level_tuning", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p,
                                                  _c_int_p]),
    ("vfs_batch_set_level_tuning", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 5),
    ("vfs_batch_time_solves", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_float_p]),
    ("vfs_batch_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
    ("vfs_server_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p]),
    ("vfs_server_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("vfs_server_add_stream", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _c_int_p]),
    ("vfs_server_submit", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_float_p, _c_float_p, _c_float_p,
                                         ctypes.c_longlong]),
    ("vfs_server_submit_pinned", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_float_p, _c_float_p,
                                                ctypes.c_longlong]),
    ("vfs_host_alloc", ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    ("vfs_host_free", ctypes.c_int, [ctypes.c_void_p]),
    ("vfs_server_step", ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p]),
    ("vfs_server_collect", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _c_int_p]),
    ("vfs_server_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("vfs_server_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
]


class _ServerParams(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_streams", ctypes.c_int), ("max_active", ctypes.c_int),
                ("storage", ctypes.c_int)] + [(n, ctypes.c_float) for n in (
                    "beta", "gamma", "alpha0", "alpha1", "timestep_lambda", "lambda_", "fScale", "limitRange")]


class _StreamParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("width", "height", "nLevels", "nWarpIters", "nSolverIters",
                                            "warm_level", "warm_warps", "warm_iters")] + [
        ("mask", _c_float_p), ("tv", _c_float_p), ("cv", _c_float_p)]


class _ServerResult(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_longlong), ("stream", ctypes.c_int), ("warm", ctypes.c_int), ("ops", ctypes.c_int),
                ("latency_ms", ctypes.c_double), ("nonfinite", ctypes.c_int), ("energy", ctypes.c_double)]


class _Energy(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("data", ctypes.c_double), ("residual", ctypes.c_double),
                ("nonfinite", ctypes.c_int), ("pixels", ctypes.c_int)]


class _ServerStats(ctypes.Structure):
    _fields_ = [("ticks", ctypes.c_longlong), ("entries", ctypes.c_longlong), ("hbm_used", ctypes.c_size_t),
                ("resident_streams", ctypes.c_int)]

STORAGE = {"f32": 0, "fp32": 0, "f16": 1, "fp16": 1, "bf16": 2}


class VfsError(RuntimeError):
    pass


def lib_path() -> str:
    return os.path.join(_HERE, "libvfs.so")


def lib() -> ctypes.CDLL:
    """Load libvfs.so (built in-tree by __graft_entry__.build()).  Raises if absent."""
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise VfsError(f"libvfs.so not built at {path}; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        for name, res, args in EXPORTS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(_c_float_p)


def _u8ptr(a: np.ndarray):
    return a.ctypes.data_as(_c_u8_p)


class StereoTgv:
    """stereotgv.h:14-252 over libvfs.  One instance owns one device solver."""

    def __init__(self, blockWidth: int = 32, blockHeight: int = 1, strideAlignment: int = 32):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.vfs_create_ex(ctypes.byref(h), blockWidth, blockHeight, strideAlignment)
        if rc != 0:
            raise VfsError(f"vfs_create_ex failed ({rc}): {self._L.vfs_last_error(None).decode()}")
        self._h = h
        self.width = self.height = 0

    # -- helpers
    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            msg = self._L.vfs_last_error(self._h).decode()
            raise VfsError(f"{what} failed ({rc}): {msg}")
        return rc

    def close(self):
        if getattr(self, "_h", None):
            self._L.vfs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference surface
    def initialize(self, width, height, beta, gamma, alpha0, alpha1, timestep_lambda, lambda_, nLevels, fScale,
                   nWarpIters, nSolverIters):
        self._check(self._L.vfs_initialize(self._h, width, height, beta, gamma, alpha0, alpha1, timestep_lambda,
                                           lambda_, nLevels, fScale, nWarpIters, nSolverIters), "initialize")
        self.width, self.height = width, height
        return 0

    @property
    def limitRange(self) -> float:
        v = ctypes.c_float()
        self._check(self._L.vfs_get_limit_range(self._h, ctypes.byref(v)), "get_limit_range")
        return v.value

    @limitRange.setter
    def limitRange(self, v: float):
        self._check(self._L.vfs_set_limit_range(self._h, float(v)), "set_limit_range")

    def copyMaskToDevice(self, mask: np.ndarray):
        m = np.ascontiguousarray(mask, dtype=np.float32)
        self._shape(m, (self.height, self.width))
        return self._check(self._L.vfs_copy_mask(self._h, _fptr(m), m.shape[1]), "copyMaskToDevice")

    def loadVectorFields(self, translationVector: np.ndarray, calibrationVector: np.ndarray):
        tv = np.ascontiguousarray(translationVector, dtype=np.float32)
        cv = np.ascontiguousarray(calibrationVector, dtype=np.float32)
        self._shape(tv, (self.height, self.width, 2))
        self._shape(cv, (self.height, self.width, 2))
        return self._check(self._L.vfs_load_vector_fields(self._h, _fptr(tv), 2 * self.width, _fptr(cv),
                                                          2 * self.width), "loadVectorFields")

    def copyImagesToDevice(self, i0: np.ndarray, i1: np.ndarray):
        """uint8 (H, W) -> CV_8U path, float32 (H, W) -> CV_32F path, uint8 (H, W, 3) -> CV_8UC3 path."""
        if i0.dtype == np.uint8 and i0.ndim == 3:
            a = np.ascontiguousarray(i0)
            b = np.ascontiguousarray(i1, dtype=np.uint8)
            self._shape(a, (self.height, self.width, 3))
            self._shape(b, (self.height, self.width, 3))
            return self._check(self._L.vfs_copy_images_u8c3(self._h, _u8ptr(a), _u8ptr(b), 3 * self.width),
                               "copyImagesToDevice")
        if i0.dtype == np.uint8:
            a = np.ascontiguousarray(i0)
            b = np.ascontiguousarray(i1, dtype=np.uint8)
            self._shape(a, (self.height, self.width))
            self._shape(b, (self.height, self.width))
            return self._check(self._L.vfs_copy_images_u8(self._h, _u8ptr(a), _u8ptr(b), self.width),
                               "copyImagesToDevice")
        a = np.ascontiguousarray(i0, dtype=np.float32)
        b = np.ascontiguousarray(i1, dtype=np.float32)
        self._shape(a, (self.height, self.width))
        self._shape(b, (self.height, self.width))
        return self._check(self._L.vfs_copy_images_f32(self._h, _fptr(a), _fptr(b), self.width),
                           "copyImagesToDevice")

    def solveStereoForwardMasked(self):
        return self._check(self._L.vfs_solve_masked(self._h), "solveStereoForwardMasked")

    def solveStereoForward(self):
        return self._check(self._L.vfs_solve(self._h), "solveStereoForward")

    def solve_warm(self, start_level: int = 1, warps: int = 2, iters: int = 20, masked: bool = True):
        """Warm-started re-solve from this handle's previous result (serving path)."""
        return self._check(self._L.vfs_solve_warm(self._h, 1 if masked else 0, start_level, warps, iters),
                           "solve_warm")

    def copy1DDisparityToHost(self) -> np.ndarray:
        out = np.empty((self.height, self.width), np.float32)
        self._check(self._L.vfs_copy_disparity_1d(self._h, _fptr(out), self.width), "copy1DDisparityToHost")
        return out

    def copyDisparityToHost(self) -> np.ndarray:
        out = np.empty((self.height, self.width, 2), np.float32)
        self._check(self._L.vfs_copy_flow(self._h, _fptr(out), 2 * self.width), "copyDisparityToHost")
        return out

    def copyWarpedImageToHost(self) -> np.ndarray:
        out = np.empty((self.height, self.width), np.float32)
        self._check(self._L.vfs_copy_warped_image(self._h, _fptr(out), self.width), "copyWarpedImageToHost")
        return out

    def copyDisparityVisToHost(self, flowScale: float) -> np.ndarray:
        """FlowToHSV colour coding of the flow, (h, w, 3) float32 in (B, G, R) order."""
        out = np.empty((self.height, self.width, 3), np.float32)
        self._check(self._L.vfs_copy_disparity_vis(self._h, _fptr(out), 3 * self.width, float(flowScale)),
                    "copyDisparityVisToHost")
        return out

    def copyStereoToHost(self, focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3):
        depth = np.empty((self.height, self.width), np.float32)
        X = np.empty((self.height, self.width, 3), np.float32)
        self._check(self._L.vfs_copy_stereo(self._h, _fptr(depth), self.width, _fptr(X), 3 * self.width,
                                            focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3), "copyStereoToHost")
        return depth, X

    # -- inspection / timing (no reference counterpart)
    def level_dims(self, level: int):
        w, h, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._L.vfs_get_level_dims(self._h, level, ctypes.byref(w), ctypes.byref(h), ctypes.byref(s)),
                    "get_level_dims")
        return w.value, h.value, s.value

    def read_plane(self, name: str, level: int = 0) -> np.ndarray:
        w, h, _ = self.level_dims(level)
        buf = np.empty(h * w * 4, np.float32)
        nch = self._check(self._L.vfs_read_plane(self._h, name.encode(), level, _fptr(buf)), f"read_plane({name})")
        arr = buf[: h * w * nch].reshape(h, w, nch)
        return arr[..., 0] if nch == 1 else arr

    def synchronize(self):
        return self._check(self._L.vfs_synchronize(self._h), "synchronize")

    def energy(self, level: int, warp: int) -> dict:
        """FP64 energy record of (level, warp) of the last solve (vfs_get_energy):
        dict(energy, data, residual, nonfinite, pixels)."""
        e = _Energy()
        self._check(self._L.vfs_get_energy(self._h, level, warp, ctypes.byref(e)), "get_energy")
        return dict(energy=e.energy, data=e.data, residual=e.residual, nonfinite=e.nonfinite, pixels=e.pixels)

    def set_option(self, key: str, value):
        if key == "storage" and isinstance(value, str):
            value = STORAGE[value.lower()]
        return self._check(self._L.vfs_set_option(self._h, key.encode(), int(value)), "set_option")

    def get_option(self, key: str) -> int:
        v = ctypes.c_int()
        self._check(self._L.vfs_get_option(self._h, key.encode(), ctypes.byref(v)), "get_option")
        return v.value

    # tile shapes by index: 0 .. 6 are those of the "shape" option; 8x8, 4x4, 32x32,
    # 20x20, 36x25, 27x25, 14x13, 7x7, 43x34, 34x40 and 53x25 exist for the multi-iteration
    # launches only, 53x50 for the persistent launches only
    SHAPES = ((64, 8), (64, 4), (128, 4), (32, 16), (32, 8), (16, 32), (16, 16), (8, 8), (4, 4), (32, 32),
              (20, 20), (36, 25), (27, 25), (14, 13), (7, 7), (43, 34), (34, 40), (53, 25), (53, 50))

    def level_tuning(self, level: int, masked: bool = True, with_per_launch: bool = False):
        """(tile shape (w, h), Tp prefetch[, inner iterations per launch]) used at
        `level`; runs the tuner if needed."""
        sh, pre, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._L.vfs_get_level_tuning(self._h, int(masked), level, ctypes.byref(sh), ctypes.byref(pre),
                                                 ctypes.byref(k)), "get_level_tuning")
        if with_per_launch:
            return self.SHAPES[sh.value], bool(pre.value), k.value
        return self.SHAPES[sh.value], bool(pre.value)

    HALF_WORKGROUP = 0x100
    # per_launch of a level whose cold solves run a warp's iterations in one persistent launch
    # (| 2, 4 or 8: in blocks of that many iterations, halo rings exchanged between blocks)
    PERSISTENT = 0x200

    def set_level_tuning(self, level: int, shape: int, preload: bool = False, per_launch: int = 1,
                         masked: bool = True, half_workgroup: bool = False):
        """Fix the launch choices of one level (shape index as option "shape"; 1, 2, 4 or 8
        inner iterations per launch, optionally in half workgroups, or PERSISTENT, optionally
        | 2, 4 or 8)."""
        k = int(per_launch) | (self.HALF_WORKGROUP if half_workgroup else 0)
        return self._check(self._L.vfs_set_level_tuning(self._h, int(masked), level, shape, int(preload), k),
                           "set_level_tuning")

    def time_solves(self, masked: bool = True, n: int = 1) -> float:
        ms = ctypes.c_float()
        self._check(self._L.vfs_time_solves(self._h, 1 if masked else 0, n, ctypes.byref(ms)), "time_solves")
        return ms.value

    def time_inner_iterations(self, level: int, reps: int):
        ms = ctypes.c_float()
        nl = ctypes.c_int()
        self._check(self._L.vfs_time_inner_iterations(self._h, level, reps, ctypes.byref(ms), ctypes.byref(nl)),
                    "time_inner_iterations")
        return ms.value, nl.value

    def time_copy(self, nbytes: int, reps: int) -> float:
        ms = ctypes.c_float()
        self._check(self._L.vfs_time_copy(self._h, nbytes, reps, ctypes.byref(ms)), "time_copy")
        return ms.value

    def _shape(self, a: np.ndarray, shape):
        # before initialize the library itself reports the call as out of order (and
        # reads nothing)
        if self.width and a.shape != tuple(shape):
            raise VfsError(f"expected array of shape {tuple(shape)}, got {a.shape}")


class Batch:
    """Several StereoTgv solvers solved together: every kernel launch covers all pairs."""

    def __init__(self, solvers):
        self._L = lib()
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        h = ctypes.c_void_p()
        rc = self._L.vfs_batch_create(ctypes.byref(h), arr, len(self.solvers))
        if rc != 0:
            raise VfsError(f"vfs_batch_create failed ({rc}): {self._L.vfs_batch_last_error(None).decode()}")
        self._h = h

    def _check(self, rc, what):
        if rc < 0:
            raise VfsError(f"{what} failed ({rc}): {self._L.vfs_batch_last_error(self._h).decode()}")
        return rc

    def solve(self, masked: bool = True):
        return self._check(self._L.vfs_batch_solve(self._h, 1 if masked else 0), "batch_solve")

    def solve_warm(self, start_level: int = 1, warps: int = 2, iters: int = 20, masked: bool = True):
        return self._check(self._L.vfs_batch_solve_warm(self._h, 1 if masked else 0, start_level, warps, iters),
                           "batch_solve_warm")

    def synchronize(self):
        return self._check(self._L.vfs_batch_synchronize(self._h), "batch_synchronize")

    def level_tuning(self, level: int, masked: bool = True, with_per_launch: bool = False):
        """(tile shape (w, h), Tp prefetch[, inner iterations per launch]) of the batch's
        launches at `level` (after a solve)."""
        sh, pre, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._L.vfs_batch_get_level_tuning(self._h, int(masked), level, ctypes.byref(sh),
                                                        ctypes.byref(pre), ctypes.byref(k)), "batch_get_level_tuning")
        if with_per_launch:
            return StereoTgv.SHAPES[sh.value], bool(pre.value), k.value
        return StereoTgv.SHAPES[sh.value], bool(pre.value)

    def set_level_tuning(self, level: int, shape: int, preload: bool = False, per_launch: int = 1,
                         masked: bool = True, half_workgroup: bool = False):
        """Fix the batch's launch choices of one level (as StereoTgv.set_level_tuning)."""
        k = int(per_launch) | (StereoTgv.HALF_WORKGROUP if half_workgroup else 0)
        return self._check(self._L.vfs_batch_set_level_tuning(self._h, int(masked), level, shape, int(preload), k),
                           "batch_set_level_tuning")

    def time_solves(self, masked: bool = True, n: int = 1) -> float:
        ms = ctypes.c_float()
        self._check(self._L.vfs_batch_time_solves(self._h, 1 if masked else 0, n, ctypes.byref(ms)), "batch_time")
        return ms.value

    def close(self):
        if getattr(self, "_h", None):
            self._L.vfs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server for camera streams (vfs_server_* of include/vfs.h)."""

    def __init__(self, params: dict, hbm_budget_bytes: int = 64 << 30, max_streams: int = 64, max_active: int = 64,
                 storage="f32"):
        self._L = lib()
        sp = _ServerParams()
        sp.hbm_budget_bytes = int(hbm_budget_bytes)
        sp.max_streams = max_streams
        sp.max_active = max_active
        sp.storage = STORAGE[storage] if isinstance(storage, str) else int(storage)
        for k_py, k_c in (("beta", "beta"), ("gamma", "gamma"), ("alpha0", "alpha0"), ("alpha1", "alpha1"),
                          ("timestep_lambda", "timestep_lambda"), ("lam", "lambda_"), ("fScale", "fScale"),
                          ("limitRange", "limitRange")):
            setattr(sp, k_c, float(params[k_py]))
        h = ctypes.c_void_p()
        rc = self._L.vfs_server_create(ctypes.byref(h), ctypes.byref(sp))
        if rc != 0:
            raise VfsError(f"vfs_server_create failed ({rc}): {self._L.vfs_server_last_error(None).decode()}")
        self._h = h
        self._keep = {}  # result buffers by tag
        self.shapes = {}
        self._pinned = []  # vfs_host_alloc buffers of pinned_frame()

    def _check(self, rc, what):
        if rc < 0:
            raise VfsError(f"{what} failed ({rc}): {self._L.vfs_server_last_error(self._h).decode()}")
        return rc

    def add_stream(self, mask, tv, cv, levels, warps, iters, warm_level=1, warm_warps=2, warm_iters=20) -> int:
        m = np.ascontiguousarray(mask, np.float32)
        t = np.ascontiguousarray(tv, np.float32)
        c = np.ascontiguousarray(cv, np.float32)
        H, W = m.shape
        p = _StreamParams(W, H, levels, warps, iters, warm_level, warm_warps, warm_iters, _fptr(m), _fptr(t), _fptr(c))
        sid = ctypes.c_int()
        self._check(self._L.vfs_server_add_stream(self._h, ctypes.byref(p), ctypes.byref(sid)), "add_stream")
        self.shapes[sid.value] = (H, W)
        return sid.value

    def submit(self, stream: int, i0, i1, tag: int, want_result: bool = True):
        a = np.ascontiguousarray(i0, np.float32)
        b = np.ascontiguousarray(i1, np.float32)
        out = np.empty(self.shapes[stream], np.float32) if want_result else None
        if out is not None:
            self._keep[tag] = out
        self._check(self._L.vfs_server_submit(self._h, stream, _fptr(a), _fptr(b),
                                              _fptr(out) if out is not None else None, tag), "submit")

    def pinned_frame(self, stream: int, i0, i1) -> np.ndarray:
        """A frame (i0 then i1) in page-locked host memory for submit_pinned: a float32
        array of shape (2, H, W) over a vfs_host_alloc buffer, freed with the server."""
        H, W = self.shapes[stream]
        p = ctypes.c_void_p()
        self._check(self._L.vfs_host_alloc(2 * H * W * 4, ctypes.byref(p)), "host_alloc")
        self._pinned.append(p)
        arr = np.ctypeslib.as_array((ctypes.c_float * (2 * H * W)).from_address(p.value)).reshape(2, H, W)
        arr[0] = i0
        arr[1] = i1
        return arr

    def submit_pinned(self, stream: int, frame: np.ndarray, tag: int, want_result: bool = True):
        """Queue a pinned_frame() without copying it (it must stay unchanged until collected)."""
        out = np.empty(self.shapes[stream], np.float32) if want_result else None
        if out is not None:
            self._keep[tag] = out
        self._check(self._L.vfs_server_submit_pinned(self._h, stream, _fptr(frame), _fptr(out) if out is not None
                                                     else None, tag), "submit_pinned")

    def step(self):
        act, q = ctypes.c_int(), ctypes.c_int()
        self._check(self._L.vfs_server_step(self._h, ctypes.byref(act), ctypes.byref(q)), "step")
        return act.value, q.value

    def collect(self, max_n: int = 256):
        buf = (_ServerResult * max_n)()
        n = ctypes.c_int()
        self._check(self._L.vfs_server_collect(self._h, buf, max_n, ctypes.byref(n)), "collect")
        out = []
        for r in buf[: n.value]:
            out.append(dict(tag=r.tag, stream=r.stream, warm=r.warm, ops=r.ops, latency_ms=r.latency_ms,
                            nonfinite=r.nonfinite, energy=r.energy, u=self._keep.pop(r.tag, None)))
        return out

    def stats(self):
        st = _ServerStats()
        self._check(self._L.vfs_server_stats(self._h, ctypes.byref(st)), "stats")
        return dict(ticks=st.ticks, entries=st.entries, hbm_used=st.hbm_used, resident_streams=st.resident_streams)

    def run_until_idle(self, timeout_s: float = 120.0):
        import time
        t0 = time.perf_counter()
        results = []
        while True:
            act, q = self.step()
            results += self.collect()
            if act == 0 and q == 0:
                results += self.collect()
                return results
            if time.perf_counter() - t0 > timeout_s:
                raise VfsError("server did not drain in time")

    def close(self):
        if getattr(self, "_h", None):
            self._L.vfs_server_destroy(self._h)
            self._h = None
        for p in getattr(self, "_pinned", []):
            self._L.vfs_host_free(p)
        self._pinned = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
