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


def _float_gate(w, h):
    """Outside-circle pixels of the unmasked kernels in float arithmetic:
    sqrtf((y - h/2)^2 + (x - w/2)^2) >= w / 2.2 (updateDualVariables.cu:13-22), with
    numpy's correctly rounded float32 square root."""
    f = np.float32
    ys, xs = np.mgrid[0:h, 0:w]
    dy = ys.astype(f) - f(h) / f(2)
    dx = xs.astype(f) - f(w) / f(2)
    return np.sqrt(dy * dy + dx * dx) >= f(w) / f(2.2)


def test_unmasked_circle_gate_rounds_in_float(synth):
    """At 523x392 eight pixels lie where the float square root rounds up onto the bound
    while the exact radius is below it: the reference (float sqrtf) treats them as
    outside.  The oracle's unmasked solve writes u = 0 at exactly the float gate's
    outside pixels, those eight included."""
    W, H = 523, 392
    out = _float_gate(W, H)
    ys, xs = np.mgrid[0:H, 0:W]
    exact = (xs - W / 2) ** 2 + (ys - H / 2) ** 2 >= float(np.float32(W) / np.float32(2.2)) ** 2
    assert (out != exact).sum() == 8
    pair = synth.make_pair(W, H, seed=8, radius=H / 2, dmax=6.0)
    u = O.solve_pair(pair, 1, 1, 2, synth.SOLVER_PARAMS, masked=False).copy1DDisparityToHost()
    assert np.array_equal(u == 0, out)


# ------------------------------------------------------------- pins added for the warp
# direction resample and the primal step's tensor term

def _mirror(k, n):
    """Index on the symmetric extension T[-1-i] = T[i], T[n+i] = T[n-1-i]."""
    k = np.mod(k, 2 * n)
    return np.where(k < n, k, 2 * n - 1 - k)


def test_tv2_resample_follows_the_flow():
    """FindWarpingVector (warpImage.cu:96-111): tv2(x) = tv sampled at x + warpUV(x).
    Zero flow is the identity; a per-pixel integer flow picks the texel it points at,
    mirrored at the edges; a half-pixel flow averages the neighbours; a constant field is
    invariant under any flow; the two channels are sampled independently."""
    H, W = 9, 11
    tv = rng.normal(size=(H, W, 2))
    assert np.array_equal(O.k_find_tv2(np.zeros((H, W, 2)), tv), tv)
    dx = rng.integers(-4, 5, size=(H, W))
    dy = rng.integers(-4, 5, size=(H, W))
    uv = np.stack([dx, dy], -1).astype(np.float64)
    ys, xs = np.mgrid[0:H, 0:W]
    want = tv[_mirror(ys + dy, H), _mirror(xs + dx, W)]
    assert np.allclose(O.k_find_tv2(uv, tv), want, rtol=0, atol=1e-14)
    # channel independence: a field that is zero in channel 1 stays zero there
    tv0 = tv.copy()
    tv0[..., 1] = 0.0
    got = O.k_find_tv2(uv, tv0)
    assert not got[..., 1].any() and np.allclose(got[..., 0], want[..., 0], rtol=0, atol=1e-14)
    # half a pixel right and a quarter down at interior pixels
    uv = np.zeros((H, W, 2))
    uv[..., 0] = 0.5
    uv[..., 1] = 0.25
    got = O.k_find_tv2(uv, tv)
    top = 0.5 * (tv[:-1, :-1] + tv[:-1, 1:])
    bot = 0.5 * (tv[1:, :-1] + tv[1:, 1:])
    assert np.allclose(got[:-1, :-1], 0.75 * top + 0.25 * bot, rtol=0, atol=1e-14)
    # the last column mirrors onto itself: (w-1) + 0.5 averages texel w-1 with its mirror
    assert np.allclose(got[:-1, -1], 0.75 * tv[:-1, -1] + 0.25 * tv[1:, -1], rtol=0, atol=1e-14)
    const = np.empty((H, W, 2))
    const[..., 0], const[..., 1] = 0.3, -0.7
    assert np.allclose(O.k_find_tv2(rng.normal(size=(H, W, 2)) * 5, const), const, rtol=0, atol=1e-14)


def test_primal_tensor_term_along_eigenvectors():
    """The v step's alpha1 T p term (updatePrimalVariables.cu:216-226): with q = 0 and
    v_ = 0 the step is v = tau alpha1 (T p) / eta_v per component.  T is assembled from
    its eigendecomposition T = w n n^T + t t^T, t perpendicular to n (the structure
    calcTensor.cu:60-70 builds), so T n = w n and T t = t: for p = alpha n + beta t the
    expected T p = alpha w n + beta t needs no product with a, b, c."""
    H, W = 6, 7
    mask = _full((H, W))
    th = rng.uniform(0, 2 * np.pi, (H, W))
    w = rng.uniform(0.05, 0.95, (H, W))
    nx, ny = np.cos(th), np.sin(th)
    tx, ty = -ny, nx
    a = w * nx * nx + tx * tx
    b = w * ny * ny + ty * ty
    c = w * nx * ny + tx * ty
    al = rng.normal(size=(H, W))
    be = rng.normal(size=(H, W))
    p = np.stack([al * nx + be * tx, al * ny + be * ty], -1)
    e1 = rng.uniform(2.0, 6.0, (H, W))
    e2 = rng.uniform(2.0, 6.0, (H, W))
    tau, alpha0, alpha1 = 0.7, 17.0, 1.2
    zeros = np.zeros((H, W))
    v, _, _ = O.k_primal_masked(mask, zeros, np.zeros((H, W, 2)), p, np.zeros((H, W, 4)), a, b, c, tau, e1, e2,
                                alpha0, alpha1, 0.5, zeros)
    tpx = al * w * nx + be * tx
    tpy = al * w * ny + be * ty
    assert np.allclose(v[..., 0], tau * alpha1 * tpx / e1, rtol=1e-12, atol=1e-14)
    assert np.allclose(v[..., 1], tau * alpha1 * tpy / e2, rtol=1e-12, atol=1e-14)
    # pure eigenvectors: p = n is scaled by w, p = t is kept
    vn, _, _ = O.k_primal_masked(mask, zeros, np.zeros((H, W, 2)), np.stack([nx, ny], -1), np.zeros((H, W, 4)), a, b,
                                 c, tau, e1, e1, alpha0, alpha1, 0.5, zeros)
    assert np.allclose(vn * e1[..., None] / (tau * alpha1), w[..., None] * np.stack([nx, ny], -1), rtol=1e-12,
                       atol=1e-14)
    vt, _, _ = O.k_primal_masked(mask, zeros, np.zeros((H, W, 2)), np.stack([tx, ty], -1), np.zeros((H, W, 4)), a, b,
                                 c, tau, e2, e2, alpha0, alpha1, 0.5, zeros)
    assert np.allclose(vt * e2[..., None] / (tau * alpha1), np.stack([tx, ty], -1), rtol=1e-12, atol=1e-14)


# ------------------------------------------------------------- energy monitor

def _energy(u, v, Iu, Iz, a, b, c, uprev=None, ulast=None, mask=None, lam=5.0, a0=17.0, a1=1.2):
    uprev = u if uprev is None else uprev
    ulast = u if ulast is None else ulast
    return O.k_energy(u, uprev, v, Iu, Iz, ulast, a, b, c, lam, a0, a1, mask)


def test_energy_data_term_alone():
    """Constant u equal to u_last and v = 0: no gradient anywhere (forward and backward
    differences vanish), so the energy is the data term lambda sum |Iz|."""
    H, W = 7, 9
    u = np.full((H, W), 0.37)
    Iu, Iz = rng.normal(size=(H, W)), rng.normal(size=(H, W))
    a, b, c = rng.random((H, W)), rng.random((H, W)), rng.random((H, W))
    e = _energy(u, np.zeros((H, W, 2)), Iu, Iz, a, b, c, mask=_full((H, W)))
    want = 5.0 * np.abs(Iz).sum()
    assert np.isclose(e["energy"], want, rtol=1e-13) and np.isclose(e["data"], want, rtol=1e-13)
    assert e["residual"] == 0 and e["nonfinite"] == 0 and e["pixels"] == H * W


def test_energy_linear_u_with_matching_v_has_no_smoothness_cost():
    """u = alpha x + beta y and v = (alpha, beta): grad u - v = 0 with the forward and the
    backward (edge) differences alike, and grad v = 0; with Iu = 0 only lambda |Iz| is left."""
    H, W = 8, 10
    ys, xs = np.mgrid[0:H, 0:W].astype(float)
    al, be = 0.7, -1.3
    u = al * xs + be * ys
    v = np.stack([np.full((H, W), al), np.full((H, W), be)], -1)
    Iz = rng.normal(size=(H, W))
    e = _energy(u, v, np.zeros((H, W)), Iz, rng.random((H, W)), rng.random((H, W)), rng.random((H, W)),
                ulast=rng.normal(size=(H, W)), mask=_full((H, W)))
    assert np.isclose(e["energy"], 5.0 * np.abs(Iz).sum(), rtol=1e-12)


def test_energy_tensor_term_along_eigenvectors():
    """u a ramp k (n . x) along a fixed unit n, v = 0, T = w n n^T + t t^T: grad u = k n
    exactly (also at the edges), so |T (grad u - v)| = w |k| at every pixel."""
    H, W = 6, 8
    th = 0.6
    nx, ny = np.cos(th), np.sin(th)
    tx, ty = -ny, nx
    w = rng.uniform(0.1, 0.9, (H, W))
    a, b, c = w * nx * nx + tx * tx, w * ny * ny + ty * ty, w * nx * ny + tx * ty
    ys, xs = np.mgrid[0:H, 0:W].astype(float)
    k = 1.7
    u = k * (nx * xs + ny * ys)
    e = _energy(u, np.zeros((H, W, 2)), np.zeros((H, W)), np.zeros((H, W)), a, b, c, mask=_full((H, W)))
    assert e["data"] == 0
    assert np.isclose(e["energy"], 1.2 * k * w.sum(), rtol=1e-12)


def test_energy_second_order_term_and_residual():
    """T = 0 and v = (gamma x, 0): grad v = (gamma, 0, 0, 0) everywhere (backward at the
    right edge), so the energy is alpha0 N |gamma|; the residual is sum |u - u_prev|."""
    H, W = 5, 7
    ys, xs = np.mgrid[0:H, 0:W].astype(float)
    g = -0.45
    v = np.stack([g * xs, np.zeros((H, W))], -1)
    u = rng.normal(size=(H, W))
    du = rng.normal(size=(H, W))
    z = np.zeros((H, W))
    e = _energy(u, v, z, z, z, z, z, uprev=u - du, ulast=u, mask=_full((H, W)))
    assert np.isclose(e["energy"], 17.0 * H * W * abs(g), rtol=1e-12)
    assert np.isclose(e["residual"], np.abs(du).sum(), rtol=1e-12)


def test_energy_backward_difference_at_mask_edge_and_write_set():
    """u = x^2, T = I, v = 0: |u_x| is the forward difference 2x + 1 inside, the backward
    one 2x - 1 where the right neighbour is outside the mask or the image; only mask pixels
    are counted."""
    H, W = 4, 9
    ys, xs = np.mgrid[0:H, 0:W].astype(float)
    u = xs ** 2
    mask = _full((H, W))
    mask[:, 5] = 0.0
    one, z = np.ones((H, W)), np.zeros((H, W))
    e = _energy(u, np.zeros((H, W, 2)), z, z, one, one, z, mask=mask)
    per_col = [2 * x + 1 if (x + 1 < W and x + 1 != 5) else 2 * x - 1 for x in range(W) if x != 5]
    assert e["pixels"] == H * (W - 1)
    assert np.isclose(e["energy"], 1.2 * H * sum(per_col), rtol=1e-12)


def test_energy_counts_nonfinite_pixels():
    """A NaN u at an interior pixel makes its own term and those of its left and upper
    neighbours (whose forward differences reach it) non-finite: 3 pixels, left out of the
    sums."""
    H, W = 6, 6
    u = rng.normal(size=(H, W))
    u[3, 3] = np.nan
    z = np.zeros((H, W))
    one = np.ones((H, W))
    e = _energy(u, np.zeros((H, W, 2)), z, z, one, one, z, uprev=np.nan_to_num(u), mask=_full((H, W)))
    assert e["nonfinite"] == 3 and np.isfinite(e["energy"]) and e["pixels"] == H * W


def test_energy_unmasked_counts_the_circle():
    H, W = 40, 44
    z = np.zeros((H, W))
    e = _energy(z, np.zeros((H, W, 2)), z, z, z, z, z)
    assert e["pixels"] == int((~_float_gate(W, H)).sum())


def test_energy_records_of_a_solve(synth):
    """The oracle's solve records every (level, warp): the data term is the data part of
    the energy, every mask pixel of the level is counted, and nothing is non-finite."""
    pair = synth.make_pair(96, 80, seed=17, radius=40, dmax=4.0)
    o = O.solve_pair(pair, 2, 2, 6, synth.SOLVER_PARAMS)
    for lvl in range(2):
        npx = int((o.read_plane("mask", lvl) != 0).sum())
        for wi in range(2):
            r = o.energy(lvl, wi)
            assert r["pixels"] == npx and r["nonfinite"] == 0
            assert 0 < r["data"] < r["energy"] and r["residual"] > 0
