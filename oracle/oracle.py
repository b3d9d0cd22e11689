"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path (vfs_amd/) never
imports it.

`Oracle(precision="f64")` mirrors the StereoTgv class surface (stereotgv.h:107-124)
with the FP64 restatement; precision="f32" runs the same algorithm in float.
The `k_*` functions expose single reference kernels on compact float64 arrays for the
closed-form / invariant tests.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    """Compile liboracle.so from the C restatement (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
        for sfx in ("f64", "f32"):
            getattr(_LIB, f"oracle_create_{sfx}").restype = ctypes.c_void_p
            getattr(_LIB, f"oracle_create_{sfx}").argtypes = (
                [ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 6 + [ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                                                       ctypes.c_int, ctypes.c_int, ctypes.c_int])
            for fn in ("oracle_destroy", "oracle_solve", "oracle_solve_warm", "oracle_copy_mask", "oracle_load_vector_fields",
                       "oracle_copy_images_f32", "oracle_copy_images_u8", "oracle_copy_images_u8c3",
                       "oracle_set_limit_range", "oracle_get_u", "oracle_get_warpuv", "oracle_get_i1warp",
                       "oracle_read_plane", "oracle_level_dims", "oracle_stereo"):
                getattr(_LIB, f"{fn}_{sfx}").restype = ctypes.c_int
    return _LIB


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's loops."""
    lib().oracle_set_threads(int(n))


def get_threads() -> int:
    lib().oracle_get_threads.restype = ctypes.c_int
    return int(lib().oracle_get_threads())


def _p(a: np.ndarray, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))


class Oracle:
    def __init__(self, width, height, beta, gamma, alpha0, alpha1, timestep_lambda, lambda_, nLevels, fScale,
                 nWarpIters, nSolverIters, precision="f64", blockWidth=32, strideAlignment=32):
        self.L = lib()
        self.sfx = precision
        self.dt = np.float64 if precision == "f64" else np.float32
        self.ct = ctypes.c_double if precision == "f64" else ctypes.c_float
        self.h = getattr(self.L, f"oracle_create_{self.sfx}")(
            width, height, beta, gamma, alpha0, alpha1, timestep_lambda, lambda_, nLevels, fScale, nWarpIters,
            nSolverIters, blockWidth, strideAlignment)
        if not self.h:
            raise ValueError("oracle_create rejected the parameters")
        self.width, self.height = width, height
        self.nLevels = nLevels

    def _fn(self, name):
        return getattr(self.L, f"{name}_{self.sfx}")

    def close(self):
        if self.h:
            self._fn("oracle_destroy")(ctypes.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def _h(self):
        return ctypes.c_void_p(self.h)

    def set_limit_range(self, v):
        self._fn("oracle_set_limit_range")(self._h, ctypes.c_float(v))

    def copyMaskToDevice(self, mask):
        m = np.ascontiguousarray(mask, np.float32)
        self._fn("oracle_copy_mask")(self._h, _p(m, ctypes.c_float), m.shape[1])

    def loadVectorFields(self, tv, cv):
        t = np.ascontiguousarray(tv, np.float32)
        c = np.ascontiguousarray(cv, np.float32)
        self._fn("oracle_load_vector_fields")(self._h, _p(t, ctypes.c_float), 2 * self.width,
                                              _p(c, ctypes.c_float), 2 * self.width)

    def copyImagesToDevice(self, i0, i1):
        if i0.dtype == np.uint8 and i0.ndim == 3:
            a, b = np.ascontiguousarray(i0), np.ascontiguousarray(i1)
            self._fn("oracle_copy_images_u8c3")(self._h, _p(a, ctypes.c_uint8), _p(b, ctypes.c_uint8), 3 * self.width)
        elif i0.dtype == np.uint8:
            a, b = np.ascontiguousarray(i0), np.ascontiguousarray(i1)
            self._fn("oracle_copy_images_u8")(self._h, _p(a, ctypes.c_uint8), _p(b, ctypes.c_uint8), self.width)
        else:
            a = np.ascontiguousarray(i0, np.float32)
            b = np.ascontiguousarray(i1, np.float32)
            self._fn("oracle_copy_images_f32")(self._h, _p(a, ctypes.c_float), _p(b, ctypes.c_float), self.width)

    def solveStereoForwardMasked(self):
        return self._fn("oracle_solve")(self._h, 1)

    def solveStereoForward(self):
        return self._fn("oracle_solve")(self._h, 0)

    def solve_warm(self, start_level=1, warps=2, iters=20, masked=True):
        return self._fn("oracle_solve_warm")(self._h, 1 if masked else 0, start_level, warps, iters)

    def copy1DDisparityToHost(self):
        out = np.empty((self.height, self.width), self.dt)
        self._fn("oracle_get_u")(self._h, _p(out, self.ct))
        return out

    def copyDisparityToHost(self):
        out = np.empty((self.height, self.width, 2), self.dt)
        self._fn("oracle_get_warpuv")(self._h, _p(out, self.ct))
        return out

    def copyWarpedImageToHost(self):
        out = np.empty((self.height, self.width), self.dt)
        self._fn("oracle_get_i1warp")(self._h, _p(out, self.ct))
        return out

    def copyStereoToHost(self, focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3):
        depth = np.empty((self.height, self.width), self.dt)
        X = np.empty((self.height, self.width, 3), self.dt)
        args = [ctypes.c_float(v) for v in (focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3)]
        self._fn("oracle_stereo")(self._h, *args, _p(depth, self.ct), _p(X, self.ct))
        return depth, X

    def energy(self, level, warp):
        """Energy record of (level, warp) of the last solve: dict(energy, data, residual,
        nonfinite, pixels) (oracle_k_energy)."""
        out = np.zeros(5)
        f = self._fn("oracle_get_energy")
        f.restype = None
        f(self._h, level, warp, _p(out))
        return dict(energy=out[0], data=out[1], residual=out[2], nonfinite=int(out[3]), pixels=int(out[4]))

    def level_dims(self, level):
        w, h, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._fn("oracle_level_dims")(self._h, level, ctypes.byref(w), ctypes.byref(h), ctypes.byref(s))
        return w.value, h.value, s.value

    def read_plane(self, name, level=0):
        w, h, _ = self.level_dims(level)
        buf = np.empty(h * w * 4, self.dt)
        nch = self._fn("oracle_read_plane")(self._h, name.encode(), level, _p(buf, self.ct))
        if nch <= 0:
            raise KeyError(name)
        arr = buf[: h * w * nch].reshape(h, w, nch)
        return arr[..., 0] if nch == 1 else arr


def solve_pair(pair: dict, levels: int, warps: int, iters: int, params: dict, precision="f64", masked=True,
               limit=None):
    """Run the oracle end to end on a synthetic pair (vfs_amd.synth.make_pair layout)."""
    H, W = pair["i0"].shape
    o = Oracle(W, H, params["beta"], params["gamma"], params["alpha0"], params["alpha1"],
               params["timestep_lambda"], params["lam"], levels, params["fScale"], warps, iters, precision)
    o.set_limit_range(params["limitRange"] if limit is None else limit)
    o.copyMaskToDevice(pair["mask"])
    o.loadVectorFields(pair["tv"], pair["cv"])
    o.copyImagesToDevice(pair["i0"], pair["i1"])
    if masked:
        o.solveStereoForwardMasked()
    else:
        o.solveStereoForward()
    return o


# ---------------------------------------------------------------- single kernels (f64)

def _d(a):
    return np.ascontiguousarray(a, np.float64)


def _dims(a):
    h, w = a.shape[:2]
    return w, h, w


def _k(name):
    f = getattr(lib(), name + "_f64")
    f.restype = None
    return f


def k_bilinear(img, px, py):
    """Bilinear+mirror fetch through the oracle's warp kernel (one output per position)."""
    img = _d(img)
    h, w = img.shape
    px = np.asarray(px, np.float64)
    py = np.asarray(py, np.float64)
    out = np.empty(px.shape)
    flat_x, flat_y, flat_o = px.ravel(), py.ravel(), out.ravel()
    # warp a 1-pixel "image" trick is awkward; evaluate row by row with a dedicated flow field
    ys, xs = np.mgrid[0:h, 0:w]
    for k in range(flat_x.size):
        uv = np.zeros((h, w, 2))
        uv[0, 0] = (flat_x[k], flat_y[k])
        res = np.empty((h, w))
        _k("oracle_k_warp")(_p(img), w, h, w, _p(uv), _p(res))
        flat_o[k] = res[0, 0]
    return out


def k_warp(img, uv):
    img, uv = _d(img), _d(uv)
    h, w = img.shape
    out = np.zeros((h, w))
    _k("oracle_k_warp")(_p(img), w, h, w, _p(uv), _p(out))
    return out


def k_find_tv2(uv, tv):
    """FindWarpingVector: the 2-channel trajectory field `tv` sampled (bilinear, mirror)
    at x + warpUV(x) for every pixel."""
    uv, tv = _d(uv), _d(tv)
    h, w = tv.shape[:2]
    out = np.zeros((h, w, 2))
    _k("oracle_k_find_tv2")(_p(uv), _p(tv), w, h, w, _p(out))
    return out


def k_downscale(src, w, h):
    src = _d(src)
    H, W = src.shape[:2]
    nch = 1 if src.ndim == 2 else src.shape[2]
    out = np.zeros((h, w) + (() if nch == 1 else (nch,)))
    _k("oracle_k_downscale")(_p(src), nch, W, H, W, w, h, w, _p(out))
    return out


def k_downscale_nn(src, w, h):
    src = _d(src)
    H, W = src.shape
    out = np.zeros((h, w))
    _k("oracle_k_downscale_nn")(_p(src), W, H, W, w, h, w, _p(out))
    return out


def k_upscale(src, W, H, scale):
    src = _d(src)
    h, w = src.shape[:2]
    nch = 1 if src.ndim == 2 else src.shape[2]
    out = np.zeros((H, W) + (() if nch == 1 else (nch,)))
    _k("oracle_k_upscale")(_p(src), nch, w, h, w, W, H, W, ctypes.c_double(scale), _p(out))
    return out


def k_gauss(img, mask=None):
    img = _d(img)
    h, w = img.shape
    out = np.zeros((h, w))
    m = None if mask is None else _d(mask)
    _k("oracle_k_gauss")(_p(img), None if m is None else _p(m), w, h, w, _p(out))
    return out


def k_tensor(gray, beta, gamma, mask=None, xlimit=None):
    gray = _d(gray)
    h, w = gray.shape
    a, b, c = np.zeros((h, w)), np.zeros((h, w)), np.zeros((h, w))
    m = None if mask is None else _d(mask)
    xl = w if xlimit is None else xlimit
    _k("oracle_k_tensor")(_p(gray), None if m is None else _p(m), ctypes.c_double(beta), ctypes.c_double(gamma),
                          w, h, w, xl, _p(a), _p(b), _p(c))
    return a, b, c


def k_eta(a, b, c, alpha0, alpha1, mask=None):
    a, b, c = _d(a), _d(b), _d(c)
    h, w = a.shape
    eu, e1, e2 = np.zeros((h, w)), np.zeros((h, w)), np.zeros((h, w))
    m = None if mask is None else _d(mask)
    _k("oracle_k_eta")(None if m is None else _p(m), ctypes.c_double(alpha0), ctypes.c_double(alpha1), _p(a), _p(b),
                       _p(c), w, h, w, _p(eu), _p(e1), _p(e2))
    return eu, e1, e2


def k_deriv(I0, I1w, tv):
    I0, I1w, tv = _d(I0), _d(I1w), _d(tv)
    h, w = I0.shape
    Iu, Iz = np.zeros((h, w)), np.zeros((h, w))
    _k("oracle_deriv")(_p(I0), _p(I1w), _p(tv), w, h, w, _p(Iu), _p(Iz))
    return Iu, Iz


def k_energy(u, uprev, v, Iu, Iz, ulast, a, b, c, lam, alpha0, alpha1, mask=None):
    """Energy monitor terms (oracle_k_energy) of compact (h, w) planes; v is (h, w, 2);
    mask=None is the unmasked form (circle gate)."""
    u, uprev, v, Iu, Iz, ulast, a, b, c = map(_d, (u, uprev, v, Iu, Iz, ulast, a, b, c))
    h, w = u.shape
    out = np.zeros(5)
    m = None if mask is None else _d(mask)
    dd = ctypes.c_double
    _k("oracle_k_energy")(None if m is None else _p(m), _p(u), _p(uprev), _p(v), _p(Iu), _p(Iz), _p(ulast), _p(a),
                          _p(b), _p(c), w, h, w, dd(lam), dd(alpha0), dd(alpha1), _p(out))
    return dict(energy=out[0], data=out[1], residual=out[2], nonfinite=int(out[3]), pixels=int(out[4]))


def k_median5(u):
    u = _d(u)
    h, w = u.shape
    out = np.zeros((h, w))
    _k("oracle_k_median5")(_p(u), w, h, w, _p(out))
    return out


def k_dual_masked(mask, u_, v_, alpha0, alpha1, sigma, a, b, c, p, q, eta_p=3.0, eta_q=2.0):
    mask, u_, v_, a, b, c = map(_d, (mask, u_, v_, a, b, c))
    p, q = _d(p).copy(), _d(q).copy()
    h, w = mask.shape
    gradv = np.zeros((h, w, 4))
    dd = ctypes.c_double
    _k("oracle_k_dual_masked")(_p(mask), _p(u_), _p(v_), dd(alpha0), dd(alpha1), dd(sigma), dd(eta_p), dd(eta_q),
                               _p(a), _p(b), _p(c), w, h, w, _p(gradv), _p(p), _p(q))
    return p, q, gradv


def k_tp(a, b, c, p, mask=None):
    a, b, c, p = map(_d, (a, b, c, p))
    h, w = a.shape
    Tp = np.zeros((h, w, 2))
    m = None if mask is None else _d(mask)
    _k("oracle_k_tp")(None if m is None else _p(m), _p(a), _p(b), _p(c), _p(p), w, h, w, _p(Tp))
    return Tp


def k_thresh_masked(Tp, u_, Iu, Iz, mask, lam, tau, etau, u):
    Tp, u_, Iu, Iz, mask, etau, u = map(_d, (Tp, u_, Iu, Iz, mask, etau, u))
    h, w = u.shape
    us = np.zeros((h, w))
    dd = ctypes.c_double
    _k("oracle_k_thresh_masked")(_p(Tp), _p(u_), _p(Iu), _p(Iz), _p(mask), dd(lam), dd(tau), _p(etau), _p(u),
                                 _p(us), w, h, w)
    return us


def k_primal_masked(mask, u_, v_, p, q, a, b, c, tau, etav1, etav2, alpha0, alpha1, mu, u):
    mask, u_, v_, p, q, a, b, c, etav1, etav2, u = map(_d, (mask, u_, v_, p, q, a, b, c, etav1, etav2, u))
    h, w = mask.shape
    v, u_s, v_s = np.zeros((h, w, 2)), np.zeros((h, w)), np.zeros((h, w, 2))
    dd = ctypes.c_double
    _k("oracle_k_primal_masked")(_p(mask), _p(u_), _p(v_), _p(p), _p(q), _p(a), _p(b), _p(c), dd(tau), _p(etav1),
                                 _p(etav2), dd(alpha0), dd(alpha1), dd(mu), _p(u), _p(v), _p(u_s), _p(v_s), w, h, w)
    return v, u_s, v_s


def k_flow_hsv(uv, scale):
    """FlowToHSV colour coding of a (h, w, 2) flow: (h, w, 3) in (B, G, R) order."""
    uv = _d(uv)
    h, w = uv.shape[:2]
    out = np.zeros((h, w, 3))
    _k("oracle_k_flow_hsv")(_p(uv), ctypes.c_double(scale), _p(out), w, h, w)
    return out


def wheel_entry(k):
    """Entry k (0..54) of the oracle's colour wheel as (R, G, B) in 0..255."""
    out = np.zeros(3)
    _k("oracle_wheel")(ctypes.c_int(k), _p(out))
    return out


def k_convert_kb(uv, focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3):
    uv = _d(uv)
    h, w = uv.shape[:2]
    X, depth = np.zeros((h, w, 3)), np.zeros((h, w))
    dd = ctypes.c_double
    _k("oracle_k_convert_kb")(_p(uv), *[dd(v) for v in (focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3)],
                              _p(X), _p(depth), w, h, w)
    return depth, X
