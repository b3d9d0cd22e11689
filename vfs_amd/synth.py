"""Seeded synthetic fisheye-stereo pairs for the StereoTgv solver.

The reference ships no runnable full-resolution inputs (its `.flo` trajectory /
calibration fields at 848x800 are absent), so workloads are generated:

* I0: 3-octave value noise (lattice 16/8/4 px, amplitudes 1, 1/2, 1/4, bilinear
  lattice interpolation), normalised to [0, 1].
* mask: pixel centres inside the circle of radius R around (W/2, H/2), as float 0/1
  (the form test_vfs/main.cpp:189 hands to copyMaskToDevice).
* trajectory tv: unit vectors radial from (W/2, H/2) through the pixel centre.
* calibration cv: i.i.d. N(0, 0.3^2) px per channel.
* ground-truth disparity: 8 seeded Voronoi cells, each a plane
  d0 + gx (x-cx)/W + gy (y-cy)/H (d0 in U[.2,.8] Dmax, g in U[-.3,.3] Dmax), clamped
  to [0, Dmax], then Gaussian-blurred (sigma 3 px).
* I1(y) = bilinear(I0, y - d(y) t(y)) + N(0, 0.01^2), so that warping I1 by
  u * tv (warpImage.cu:9-23) with u = d maps it back onto I0.
"""
from __future__ import annotations

import numpy as np

# Solver settings shared by every workload (test_vfs/main.cpp:141-146 values,
# limitRange at the class default 1.0, stereotgv.h:45).
SOLVER_PARAMS = dict(beta=9.0, gamma=0.85, alpha0=17.0, alpha1=1.2, timestep_lambda=1.0,
                     lam=5.0, fScale=2.0, limitRange=1.0)

# Named workloads: (W, H, levels, warps, iterations, mask radius, Dmax, seed)
WORKLOADS = {
    "tgv_424x400": dict(W=424, H=400, levels=4, warps=5, iters=20, radius=200, dmax=20.0, seed=1),
    "tgv_848x800": dict(W=848, H=800, levels=5, warps=10, iters=50, radius=400, dmax=40.0, seed=2),
    "tgv_640x480": dict(W=640, H=480, levels=5, warps=10, iters=50, radius=240, dmax=30.0, seed=100),
}


def _mirror(k: np.ndarray, n: int) -> np.ndarray:
    period = 2 * n
    m = np.mod(k, period)
    return np.where(m < n, m, period - 1 - m)


def bilinear_mirror(img: np.ndarray, px: np.ndarray, py: np.ndarray) -> np.ndarray:
    """Bilinear sample of a 2-D array at pixel-index positions, mirror addressing."""
    h, w = img.shape
    fx = np.floor(px)
    fy = np.floor(py)
    ax = px - fx
    ay = py - fy
    i = fx.astype(np.int64)
    j = fy.astype(np.int64)
    i0, i1 = _mirror(i, w), _mirror(i + 1, w)
    j0, j1 = _mirror(j, h), _mirror(j + 1, h)
    top = (1 - ax) * img[j0, i0] + ax * img[j0, i1]
    bot = (1 - ax) * img[j1, i0] + ax * img[j1, i1]
    return (1 - ay) * top + ay * bot


def value_noise(W: int, H: int, rng: np.random.Generator) -> np.ndarray:
    img = np.zeros((H, W), np.float64)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    for lattice, amp in ((16, 1.0), (8, 0.5), (4, 0.25)):
        gh, gw = H // lattice + 2, W // lattice + 2
        grid = rng.random((gh, gw))
        gx = (xs + 0.5) / lattice
        gy = (ys + 0.5) / lattice
        x0 = np.floor(gx).astype(np.int64)
        y0 = np.floor(gy).astype(np.int64)
        fx = gx - x0
        fy = gy - y0
        v = ((1 - fy) * ((1 - fx) * grid[y0, x0] + fx * grid[y0, x0 + 1])
             + fy * ((1 - fx) * grid[y0 + 1, x0] + fx * grid[y0 + 1, x0 + 1]))
        img += amp * v
    img -= img.min()
    img /= max(img.max(), 1e-12)
    return img


def _gaussian_blur(img: np.ndarray, sigma: float) -> np.ndarray:
    rad = int(np.ceil(3 * sigma))
    k = np.exp(-0.5 * (np.arange(-rad, rad + 1) / sigma) ** 2)
    k /= k.sum()
    pad = np.pad(img, rad, mode="edge")
    tmp = np.zeros_like(pad[rad:-rad, :])
    for i, kv in enumerate(k):
        tmp += kv * pad[i:i + img.shape[0], :]
    out = np.zeros_like(img)
    for i, kv in enumerate(k):
        out += kv * tmp[:, i:i + img.shape[1]]
    return out


def gt_disparity(W: int, H: int, dmax: float, rng: np.random.Generator) -> np.ndarray:
    cx, cy = W / 2.0, H / 2.0
    sites = np.stack([rng.random(8) * W, rng.random(8) * H], axis=1)
    d0 = rng.uniform(0.2, 0.8, 8) * dmax
    g = rng.uniform(-0.3, 0.3, (8, 2)) * dmax
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    dist = (xs[None] + 0.5 - sites[:, 0, None, None]) ** 2 + (ys[None] + 0.5 - sites[:, 1, None, None]) ** 2
    cell = np.argmin(dist, axis=0)
    d = d0[cell] + g[cell, 0] * (xs - cx) / W + g[cell, 1] * (ys - cy) / H
    d = np.clip(d, 0.0, dmax)
    return _gaussian_blur(d, 3.0)


def make_pair(W: int, H: int, seed: int, radius: float | None = None, dmax: float | None = None) -> dict:
    """One synthetic pair.  Returns float32 arrays:
    i0, i1 (H, W); mask (H, W) 0/1; tv, cv (H, W, 2); gt (H, W)."""
    rng = np.random.default_rng(seed)
    radius = H / 2.0 if radius is None else radius
    dmax = W / 21.2 if dmax is None else dmax
    cx, cy = W / 2.0, H / 2.0
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    i0 = value_noise(W, H, rng)
    mask = (((xs + 0.5 - cx) ** 2 + (ys + 0.5 - cy) ** 2) <= radius ** 2).astype(np.float32)
    tx = xs + 0.5 - cx
    ty = ys + 0.5 - cy
    r = np.sqrt(tx * tx + ty * ty)
    tv = np.stack([tx / r, ty / r], axis=-1)
    cv = rng.normal(0.0, 0.3, (H, W, 2))
    d = gt_disparity(W, H, dmax, rng)
    i1 = bilinear_mirror(i0, xs - d * tv[..., 0], ys - d * tv[..., 1])
    i1 = i1 + rng.normal(0.0, 0.01, (H, W))
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return dict(i0=f32(i0), i1=f32(i1), mask=f32(mask), tv=f32(tv), cv=f32(cv), gt=f32(d))


def drift_pair(base: dict, frame: int, seed: int, step: float = 0.5) -> dict:
    """Frame `frame` of a camera stream: the base disparity shifted by a smooth
    per-frame drift of at most `step` px (serving workload)."""
    phase = np.random.default_rng(seed).random(2) * 2 * np.pi  # one drift pattern per stream
    rng = np.random.default_rng(seed * 1000003 + frame)          # fresh sensor noise per frame
    H, W = base["i0"].shape
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    drift = step * frame * 0.5 * (np.sin(2 * np.pi * xs / W + phase[0]) * 0.5
                                  + np.cos(2 * np.pi * ys / H + phase[1]) * 0.5)
    d = np.clip(base["gt"].astype(np.float64) + drift, 0.0, None)
    tv = base["tv"].astype(np.float64)
    i1 = bilinear_mirror(base["i0"].astype(np.float64), xs - d * tv[..., 0], ys - d * tv[..., 1])
    i1 = i1 + rng.normal(0.0, 0.01, (H, W))
    out = dict(base)
    out["i1"] = np.ascontiguousarray(i1, dtype=np.float32)
    out["gt"] = np.ascontiguousarray(d, dtype=np.float32)
    return out
