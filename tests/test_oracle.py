"""Pins of the CPU oracle (no GPU).  The reference ships no golden vectors or
known-answer tests for the solver, so every oracle kernel is checked against what
the mathematics fixes: closed forms, invariants and brute force on tiny inputs,
each chosen so that a plausible transcription mistake (wrong tap, sign, index,
component pairing or boundary rule) fails.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

rng = np.random.default_rng(12345)


# ------------------------------------------------------------- sampling / pyramids

def test_warp_zero_flow_is_identity_and_integer_flow_is_shift():
    img = rng.random((9, 13))
    uv = np.zeros((9, 13, 2))
    assert np.array_equal(O.k_warp(img, uv), img)
    uv[..., 0] = 2.0
    uv[..., 1] = -1.0
    out = O.k_warp(img, uv)
    # interior: out(x, y) = img(x + 2, y - 1)
    assert np.allclose(out[1:, :-2], img[:-1, 2:])


def test_warp_mirror_addressing():
    img = rng.random((6, 7))
    uv = np.zeros((6, 7, 2))
    uv[:, 0, 0] = -1.0  # x = -1 mirrors to 0
    uv[:, 6, 0] = 1.0   # x = 7 mirrors to 6
    uv[0, 3, 1] = -3.0  # y = -3 mirrors to 2
    out = O.k_warp(img, uv)
    assert np.allclose(out[:, 0], img[:, 0])
    assert np.allclose(out[:, 6], img[:, 6])
    assert np.isclose(out[0, 3], img[2, 3])


def test_warp_half_pixel_is_average():
    img = rng.random((5, 8))
    uv = np.zeros((5, 8, 2))
    uv[..., 0] = 0.5
    uv[..., 1] = 0.25
    out = O.k_warp(img, uv)
    y, x = 2, 3
    top = 0.5 * (img[y, x] + img[y, x + 1])
    bot = 0.5 * (img[y + 1, x] + img[y + 1, x + 1])
    assert np.isclose(out[y, x], 0.75 * top + 0.25 * bot)


def test_downscale_exact_2x_is_box_mean():
    img = rng.random((12, 16))
    out = O.k_downscale(img, 8, 6)
    box = img.reshape(6, 2, 8, 2).mean(axis=(1, 3))
    assert np.allclose(out, box, atol=1e-14)


def test_downscale_non_exact_on_linear_ramp():
    # bilinear reproduces a linear function, so the 4-tap mean of a ramp is the ramp at
    # the coarse pixel centre mapped to fine coordinates: (x + .5) W / w - .5
    W, H, w, h = 53, 50, 26, 25
    ys, xs = np.mgrid[0:H, 0:W].astype(float)
    img = 0.3 * xs + 0.7 * ys
    out = O.k_downscale(img, w, h)
    cy, cx = np.mgrid[0:h, 0:w].astype(float)
    fx = (cx + 0.5) * W / w - 0.5
    fy = (cy + 0.5) * H / h - 0.5
    inner = (fx - 0.25 * W / w > 0) & (fx + 0.25 * W / w < W - 1) & (fy - 0.25 * H / h > 0) & (fy + 0.25 * H / h < H - 1)
    assert inner.sum() > 400
    assert np.allclose(out[inner], (0.3 * fx + 0.7 * fy)[inner], atol=1e-12)


def test_downscale_float2_channels_independent():
    f = rng.random((10, 14, 2))
    out = O.k_downscale(f, 7, 5)
    assert np.allclose(out[..., 0], O.k_downscale(f[..., 0].copy(), 7, 5))
    assert np.allclose(out[..., 1], O.k_downscale(f[..., 1].copy(), 7, 5))


def test_mask_pyramid_nearest_neighbour():
    m = (rng.random((20, 24)) > 0.5).astype(float)
    out = O.k_downscale_nn(m, 12, 10)
    assert np.array_equal(out, m[1::2, 1::2])
    # non-exact ratio: texel floor((x + .5) W / w)
    out2 = O.k_downscale_nn(m, 11, 9)
    fx = np.floor((np.arange(11) + 0.5) * 24 / 11).astype(int)
    fy = np.floor((np.arange(9) + 0.5) * 20 / 9).astype(int)
    assert np.array_equal(out2, m[np.ix_(fy, fx)])


def test_upscale_exact_2x_weights_and_scale():
    c = rng.random((4, 5))
    f = O.k_upscale(c, 10, 8, 2.0)
    # interior fine pixel (2i+1, 2j+1) sits at coarse (i + .25, j + .25)
    i, j = 1, 2
    ref = 0.75 * (0.75 * c[j, i] + 0.25 * c[j, i + 1]) + 0.25 * (0.75 * c[j + 1, i] + 0.25 * c[j + 1, i + 1])
    assert np.isclose(f[2 * j + 1, 2 * i + 1], 2.0 * ref)
    # corner replicates (mirror of -0.25)
    assert np.isclose(f[0, 0], 2.0 * c[0, 0])
    # constant field scales by `scale` exactly (fScale, not the size ratio)
    assert np.allclose(O.k_upscale(np.full((3, 3), 1.5), 7, 7, 2.0), 3.0)


# ------------------------------------------------------------------ tensor, eta

def test_gaussian_impulse_response_and_constant():
    img = np.zeros((7, 7))
    img[3, 3] = 16.0
    out = O.k_gauss(img)
    k = np.outer([1, 2, 1], [1, 2, 1]).astype(float)
    assert np.allclose(out[2:5, 2:5], k)
    assert np.isclose(out.sum(), 16.0)
    assert np.allclose(O.k_gauss(np.full((5, 6), 0.4)), 0.4)


def test_gaussian_masked_writes_only_mask():
    img = rng.random((6, 6))
    mask = np.zeros((6, 6))
    mask[2:4, 1:5] = 1
    out = O.k_gauss(img, mask)
    assert np.all(out[mask == 0] == 0)
    assert np.allclose(out[mask > 0], O.k_gauss(img)[mask > 0])


def test_tensor_constant_image():
    a, b, c = O.k_tensor(np.full((6, 8), 0.3), 9.0, 0.85)
    assert np.allclose(a, 1) and np.allclose(b, 1) and np.allclose(c, 0)


def test_tensor_invariants_and_gradient_taps():
    beta, gamma = 9.0, 0.85
    gray = rng.random((9, 11))
    a, b, c = O.k_tensor(gray, beta, gamma)
    # gradient taps at the (x+1, y+1) corner (calcTensor.cu:28-39: first difference overwritten)
    g = np.zeros((9, 11, 2))
    ext = np.pad(gray, ((0, 1), (0, 1)), mode="symmetric")
    g[..., 0] = ext[1:, 1:] - ext[1:, :-1]
    g[..., 1] = ext[1:, 1:] - ext[:-1, 1:]
    mag = np.linalg.norm(g, axis=-1)
    wt = np.maximum(np.exp(-beta * mag ** gamma), 1e-8)
    # T = w n n^T + n_perp n_perp^T: trace = w + 1, det = w, T n = w n
    assert np.allclose(a + b, wt + 1)
    assert np.allclose(a * b - c * c, wt)
    ok = mag > 1e-6  # the bottom-right corner mirrors onto itself: zero gradient, n = (1, 0)
    n = g[ok] / mag[ok][:, None]
    Tn0 = a[ok] * n[:, 0] + c[ok] * n[:, 1]
    Tn1 = c[ok] * n[:, 0] + b[ok] * n[:, 1]
    assert np.allclose(Tn0, wt[ok] * n[:, 0]) and np.allclose(Tn1, wt[ok] * n[:, 1])
    assert a[-1, -1] == 1 and b[-1, -1] == 1 and c[-1, -1] == 0


def test_tensor_column_limit_leaves_memset():
    gray = rng.random((5, 12))
    a, b, c = O.k_tensor(gray, 9.0, 0.85, xlimit=8)
    assert np.all(a[:, 8:] == 0) and np.all(b[:, 8:] == 0) and np.all(c[:, 8:] == 0)
    assert np.all(a[:, :8] != 0)


def test_eta_closed_form_special_values():
    one, zero = np.ones((2, 2)), np.zeros((2, 2))
    eu, e1, e2 = O.k_eta(one, one, zero, 17.0, 1.2)
    assert np.allclose(eu, 4 * 1.2 ** 2)          # a^2 + b^2 + (a)^2 + (b)^2
    assert np.allclose(e1, 1.2 ** 2 + 4 * 17.0 ** 2)
    assert np.allclose(e2, 1.2 ** 2 + 4 * 17.0 ** 2)
    eu, e1, e2 = O.k_eta(zero, zero, one, 17.0, 1.2)
    assert np.allclose(eu, (2 + 1 + 1) * 1.2 ** 2)  # 2c^2 + c^2 + c^2
    assert np.allclose(e1, 1.2 ** 2 + 4 * 17.0 ** 2)


# ------------------------------------------------------------------- derivatives

def test_derivative_exact_on_linear_image():
    H, W = 12, 14
    ys, xs = np.mgrid[0:H, 0:W].astype(float)
    al, be = 0.3, -0.2
    I0 = al * xs + be * ys
    I1 = I0 + 0.05
    ang = rng.random((H, W)) * 2 * np.pi
    tv = np.stack([np.cos(ang), np.sin(ang)], -1) * 0.7  # non-unit: the kernel renormalises
    Iu, Iz = O.k_deriv(I0, I1, tv)
    inner = np.zeros((H, W), bool)
    inner[3:-3, 3:-3] = True
    t = tv / np.linalg.norm(tv, axis=-1, keepdims=True)
    assert np.allclose(Iu[inner], (al * t[..., 0] + be * t[..., 1])[inner], atol=1e-12)
    assert np.allclose(Iz, 0.05)


def test_derivative_zero_trajectory_is_nan():
    I = rng.random((6, 6))
    tv = np.ones((6, 6, 2))
    tv[2, 3] = 0
    Iu, _ = O.k_deriv(I, I, tv)
    assert np.isnan(Iu[2, 3]) and np.isfinite(np.delete(Iu.ravel(), 2 * 6 + 3)).all()


# ------------------------------------------------------------------------ median

def test_median_matches_brute_force_with_corner_zeros():
    u = rng.normal(size=(7, 9)) + 3.0  # positive: corner zeros pull the median down at corners
    out = O.k_median5(u)
    H, W = u.shape
    for y, x in itertools.product(range(H), range(W)):
        vals = []
        for j in range(5):
            for i in range(5):
                r, cc = y + j - 2, x + i - 2
                rin, cin = 0 <= r < H, 0 <= cc < W
                if rin and cin:
                    vals.append(u[r, cc])
                elif rin:
                    vals.append(u[r, 0 if cc < 0 else W - 1])
                elif cin:
                    vals.append(u[0 if r < 0 else H - 1, cc])
                else:
                    vals.append(0.0)
        assert out[y, x] == np.sort(vals)[12]


# ------------------------------------------------------------------- primal-dual

def _full(shape, v=1.0):
    return np.full(shape, v)


def test_dual_projection_and_fixed_point():
    H, W = 8, 9
    mask = _full((H, W))
    a, b, c = rng.random((H, W)), rng.random((H, W)), rng.random((H, W)) * 0.1
    u_ = rng.normal(size=(H, W)) * 5
    v_ = rng.normal(size=(H, W, 2)) * 5
    p0 = rng.normal(size=(H, W, 2))
    q0 = rng.normal(size=(H, W, 4))
    p, q, _ = O.k_dual_masked(mask, u_, v_, 17.0, 1.2, 1.0, a, b, c, p0, q0)
    assert np.all(np.linalg.norm(p, axis=-1) <= 1 + 1e-12)
    assert np.all(np.linalg.norm(q, axis=-1) <= 1 + 1e-12)
    # zero u_, v_: p, q only get projected
    small_p = p0 * 0.1
    small_q = q0 * 0.1
    p2, q2, _ = O.k_dual_masked(mask, np.zeros((H, W)), np.zeros((H, W, 2)), 17.0, 1.2, 1.0, a, b, c, small_p,
                                small_q)
    assert np.allclose(p2, small_p) and np.allclose(q2, small_q)


def test_dual_linear_response_is_tensor_times_gradient():
    H, W = 6, 7
    mask = _full((H, W))
    a, b, c = rng.random((H, W)), rng.random((H, W)), rng.random((H, W)) * 0.2
    u_ = rng.normal(size=(H, W))
    v_ = rng.normal(size=(H, W, 2))
    sigma, alpha1 = 1e-4, 1.2
    p, _, _ = O.k_dual_masked(mask, u_, v_, 17.0, alpha1, sigma, a, b, c, np.zeros((H, W, 2)), np.zeros((H, W, 4)))
    ux = np.empty((H, W))
    uy = np.empty((H, W))
    ux[:, :-1] = u_[:, 1:] - u_[:, :-1]
    ux[:, -1] = u_[:, -1] - u_[:, -2]      # backward at the right edge
    uy[:-1] = u_[1:] - u_[:-1]
    uy[-1] = u_[-1] - u_[-2]
    ux -= v_[..., 0]
    uy -= v_[..., 1]
    k = alpha1 * sigma / 3.0
    assert np.allclose(p[..., 0], k * (a * ux + c * uy))
    assert np.allclose(p[..., 1], k * (c * ux + b * uy))


def test_gradient_divergence_adjoint_pairing():
    """<grad v, q> = -<v, div q> for v supported away from the border: pins which q
    component pairs with which difference of which v component, and the signs."""
    H, W = 10, 12
    mask = _full((H, W))
    v = np.zeros((H, W, 2))
    v[2:-2, 2:-2] = rng.normal(size=(H - 4, W - 4, 2))
    q = rng.normal(size=(H, W, 4))
    zeros = np.zeros((H, W))
    _, _, gradv = O.k_dual_masked(mask, zeros, v, 17.0, 1.2, 0.0, zeros, zeros, zeros, np.zeros((H, W, 2)),
                                  np.zeros((H, W, 4)))
    # primal with p = 0 and v_ = 0: v = (tau / eta_v) * alpha0 * div q
    tau, alpha0, eta = 1.0, 1.0, 1.0
    e = np.full((H, W), eta)
    vout, _, _ = O.k_primal_masked(mask, zeros, np.zeros((H, W, 2)), np.zeros((H, W, 2)), q, zeros, zeros, zeros, tau,
                                   e, e, alpha0, 1.2, 0.5, zeros)
    div = vout * eta / (tau * alpha0)
    lhs = np.sum(gradv * q)
    rhs = -np.sum(v * div)
    assert np.isclose(lhs, rhs, rtol=1e-12, atol=1e-12)


def test_primal_overrelaxation():
    H, W = 5, 6
    mask = _full((H, W))
    u = rng.normal(size=(H, W))
    u_ = rng.normal(size=(H, W))
    zeros = np.zeros((H, W))
    e = np.full((H, W), 4.0)
    _, u_s, _ = O.k_primal_masked(mask, u_, np.zeros((H, W, 2)), np.zeros((H, W, 2)), np.zeros((H, W, 4)), zeros,
                                  zeros, zeros, 0.5, e, e, 17.0, 1.2, 0.8, u)
    assert np.allclose(u_s, u + 0.8 * (u - u_))


@pytest.mark.parametrize("trial", range(200))
def test_threshold_is_prox_of_l1_data_term(trial):
    """us = argmin_x (x - uhat)^2 / (2 te) + lambda |Iu (x - u) + Iz|, uhat = u_ + te div_p."""
    r = np.random.default_rng(trial)
    H, W = 1, 3
    mask = np.array([[0.0, 1.0, 0.0]])  # isolated pixel: div_p = -Tp[left].x - Tp[up].y
    Tp = np.zeros((H, W, 2))
    Tp[0, 0, 0] = r.normal()
    u_ = r.normal(size=(H, W))
    u = r.normal(size=(H, W))
    Iu = r.normal(size=(H, W)) * (0 if trial % 17 == 0 else 1)
    Iz = r.normal(size=(H, W))
    etau = np.full((H, W), r.uniform(0.5, 3))
    lam, tau = 5.0, r.uniform(0.1, 1.0)
    us = O.k_thresh_masked(Tp, u_, Iu, Iz, mask, lam, tau, etau, u)
    te = tau / etau[0, 1]
    divp = -Tp[0, 0, 0]
    uhat = u_[0, 1] + te * divp
    f = lambda x: (x - uhat) ** 2 / (2 * te) + lam * np.abs(Iu[0, 1] * (x - u[0, 1]) + Iz[0, 1])
    xs = np.linspace(us[0, 1] - 3, us[0, 1] + 3, 60001)
    assert f(us[0, 1]) <= f(xs).min() + 1e-9
    assert us[0, 0] == 0 and us[0, 2] == 0  # masked-out pixels untouched


def test_threshold_mask_boundary_rules():
    # pixel with left in mask, right out: div_x = -Tp[left].x ; up out, down in: div_y = Tp.y
    mask = np.array([[0, 0, 0], [1, 1, 0], [1, 1, 0]], float)
    Tp = rng.normal(size=(3, 3, 2))
    zeros = np.zeros((3, 3))
    eta = np.ones((3, 3))
    Iu = zeros.copy()  # Iu = 0 branch: us = uhat
    us = O.k_thresh_masked(Tp, zeros, Iu, zeros, mask, 5.0, 1.0, eta, zeros)
    assert np.isclose(us[1, 1], -Tp[1, 0, 0] + Tp[1, 1, 1])
    # pixel (2, 0): left off-image -> maskLeft 0, right in: div_x = Tp.x; down off-image: div_y = -Tp[up].y
    assert np.isclose(us[2, 0], Tp[2, 0, 0] - Tp[1, 0, 1])


# ------------------------------------------------------------------------ output

def test_convert_kb_zero_distortion_equidistant():
    """With d1..d4 = 0 the KB polynomial is theta_d = theta, so one Newton step from 0
    gives theta = r (equidistant fisheye): x_out = tan(r) x' / r, and
    Z = (t1 - x1 t3) / (x1 - x0), X = (x0 Z, y0 Z, Z), depth = |X|."""
    H, W = 4, 6
    uv = np.zeros((H, W, 2))
    uv[..., 0] = -3.0
    uv[..., 1] = 0.5
    f, c = 100.0, 200.0  # (u - focalx) / cx: the reference's argument roles
    t1, t3 = -0.06, 0.01
    depth, X = O.k_convert_kb(uv, f, f, c, c, 0, 0, 0, 0, t1, 0.0, t3)
    ys, xs = np.mgrid[0:H, 0:W].astype(float)
    xp0, yp0 = (xs - f) / c, (ys - f) / c
    xp1, yp1 = (xs - 3.0 - f) / c, (ys + 0.5 - f) / c
    r0, r1 = np.hypot(xp0, yp0), np.hypot(xp1, yp1)
    x0, y0 = np.tan(r0) * xp0 / r0, np.tan(r0) * yp0 / r0
    x1 = np.tan(r1) * xp1 / r1
    Z = (t1 - x1 * t3) / (x1 - x0)
    assert np.allclose(X[..., 2], Z, rtol=1e-12)
    assert np.allclose(X[..., 0], x0 * Z, rtol=1e-12)
    assert np.allclose(X[..., 1], y0 * Z, rtol=1e-12)
    assert np.allclose(depth, np.linalg.norm(X, axis=-1), rtol=1e-12)


# -------------------------------------------------------------------- end to end

def test_end_to_end_recovers_synthetic_disparity(synth):
    pair = synth.make_pair(160, 120, seed=3, radius=60, dmax=8.0)
    o = O.solve_pair(pair, 3, 5, 20, synth.SOLVER_PARAMS)
    u = o.copy1DDisparityToHost()
    m = pair["mask"] > 0
    err = np.abs(u - pair["gt"])[m]
    assert np.median(err) < 0.3, np.median(err)


def test_fp32_fp64_drift_small(synth):
    pair = synth.make_pair(128, 96, seed=4, radius=48, dmax=6.0)
    u64 = O.solve_pair(pair, 3, 3, 10, synth.SOLVER_PARAMS, "f64").copy1DDisparityToHost()
    u32 = O.solve_pair(pair, 3, 3, 10, synth.SOLVER_PARAMS, "f32").copy1DDisparityToHost()
    m = pair["mask"] > 0
    assert np.abs(u64 - u32)[m].mean() < 1e-3


def test_limit_range_caps_per_warp_increment(synth):
    """limitRange clamps du from above only (limitRange.cu:15-23): with a 0.05 limit,
    one level and one warp, u <= 0.05 everywhere and negative steps pass."""
    pair = synth.make_pair(96, 80, seed=5, radius=40, dmax=5.0)
    o = O.solve_pair(pair, 1, 1, 10, synth.SOLVER_PARAMS, limit=0.05)
    u = o.copy1DDisparityToHost()
    assert u.max() <= float(np.float32(0.05))  # the limit is a float
    # swapped frames: the disparity is negative and passes unclamped below -limit
    rev = dict(pair, i0=pair["i1"], i1=pair["i0"])
    u2 = O.solve_pair(rev, 1, 1, 10, synth.SOLVER_PARAMS, limit=0.05).copy1DDisparityToHost()
    assert u2.min() < -0.05


def test_unmasked_solve_zero_outside_circle(synth):
    pair = synth.make_pair(96, 80, seed=6, radius=40, dmax=4.0)
    o = O.solve_pair(pair, 1, 1, 5, synth.SOLVER_PARAMS, masked=False)
    u = o.copy1DDisparityToHost()
    ys, xs = np.mgrid[0:80, 0:96].astype(float)
    outside = (xs - 48) ** 2 + (ys - 40) ** 2 >= (np.float32(96 / 2.2)) ** 2
    # outside the circle every iteration writes 0, u = u_last + min(0 - u_last, limit) = 0
    assert np.all(u[outside] == 0)
