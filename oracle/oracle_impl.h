/*
 * CPU restatement of the StereoTgv fisheye-stereo solver (menandro/vfs, stereotgv/).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product path is
 * checked against, and the timed CPU baseline.  Nothing in the product (vfs_amd/)
 * links or calls it.
 *
 * The file is included twice (oracle_f64.c, oracle_f32.c) with
 *   R      the arithmetic type (double for the FP64 oracle, float for drift studies)
 *   FN(n)  the exported-name mangler (n##_f64 / n##_f32).
 *
 * Semantics follow the reference kernel by kernel, with the reference's buffer model:
 * every device buffer is a flat array allocated at level-0 size (stride0 * height0),
 * indexed at each level with that level's stride, and the host-side pointer swaps are
 * reproduced in the reference order.  Each kernel writes exactly the pixels the
 * reference kernel writes (masked kernels leave mask==0 pixels untouched), so values
 * left behind by earlier iterations / coarser levels are read exactly where the
 * reference reads them.
 *
 * Texture fetches (cudaFilterModeLinear + cudaAddressModeMirror, normalized coords)
 * are restated as exact bilinear interpolation at pixel-index position
 * p = x_norm * dim - 0.5 over the symmetric extension T[-1-i] = T[i],
 * T[n+i] = T[n-1-i] (periodic with period 2n).  Exact (not 9-bit fixed-point) weights.
 * Point filtering (mask pyramid) takes the texel floor(x_norm * dim).
 *
 * Deliberate deviation (reference quirk Q4): every solve starts from zeroed work
 * buffers, which equals the reference's first solve after initialize() on zeroed
 * allocations.  Out-of-allocation reads (flat index < 0) read 0.
 */
#include <tgmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORACLE_MAXL
#define ORACLE_MAXL 16
#endif

/* ------------------------------------------------------------------ helpers */

static inline long FN(mir)(long k, long n)
{
    long P = 2 * n;
    long m = k % P;
    if (m < 0) m += P;
    return m < n ? m : P - 1 - m;
}

/* Bilinear fetch of channel ch of an nch-channel pitched image at pixel-index
 * position (px, py), mirror addressing.  stereotgv texture semantics, e.g.
 * warpImage.cu:19-22, downscale.cu:28-29, upscale.cu:15-20. */
static R FN(bil)(const R *img, int nch, int ch, int w, int h, int s, R px, R py)
{
    if (px != px || py != py) return (R)NAN;
    if (px > (R)1e8) px = (R)1e8;
    if (px < (R)-1e8) px = (R)-1e8;
    if (py > (R)1e8) py = (R)1e8;
    if (py < (R)-1e8) py = (R)-1e8;
    R fx = floor(px), fy = floor(py);
    R ax = px - fx, ay = py - fy;
    long i = (long)fx, j = (long)fy;
    long i0 = FN(mir)(i, w), i1 = FN(mir)(i + 1, w);
    long j0 = FN(mir)(j, h), j1 = FN(mir)(j + 1, h);
    const R *r0 = img + (size_t)j0 * s * nch;
    const R *r1 = img + (size_t)j1 * s * nch;
    R top = ((R)1 - ax) * r0[i0 * nch + ch] + ax * r0[i1 * nch + ch];
    R bot = ((R)1 - ax) * r1[i0 * nch + ch] + ax * r1[i1 * nch + ch];
    return ((R)1 - ay) * top + ay * bot;
}

/* Exact-texel fetch with mirror addressing (integer position). */
static inline R FN(texel)(const R *img, int w, int h, int s, long x, long y)
{
    return img[(size_t)FN(mir)(y, h) * s + FN(mir)(x, w)];
}

/* Flat read that yields 0 outside the allocation (index < 0). */
static inline R FN(rd)(const R *buf, long idx) { return idx < 0 ? (R)0 : buf[idx]; }

/* ------------------------------------------------------------- elementwise */

/* imageConverter.cu:3-16  d = u8 / 256 */
void FN(oracle_k_cv8u)(const uint8_t *in, int in_pitch, R *out, int w, int h, int s)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            out[(size_t)y * s + x] = (R)in[(size_t)y * in_pitch + x] / (R)256;
}

/* imageConverter.cu:26-42  d = (r+g+b)/3/256 */
void FN(oracle_k_cv8uc3)(const uint8_t *in, int in_pitch, R *out, int w, int h, int s)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const uint8_t *px = in + (size_t)y * in_pitch + 3 * (size_t)x;
            R g = ((R)px[0] + (R)px[1] + (R)px[2]) / (R)3;
            out[(size_t)y * s + x] = g / (R)256;
        }
}

/* warpImage.cu:9-23  out(x) = bilinear(src, x + warpUV(x)) */
void FN(oracle_k_warp)(const R *src, int w, int h, int s, const R *uv, R *out)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t pos = (size_t)y * s + x;
            out[pos] = FN(bil)(src, 1, 0, w, h, s, (R)x + uv[2 * pos], (R)y + uv[2 * pos + 1]);
        }
}

/* warpImage.cu:96-111  tv2(x) = bilinear(tv, x + warpUV(x)) (2 channels) */
void FN(oracle_k_find_tv2)(const R *uv, const R *tv, int w, int h, int s, R *tv2)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t pos = (size_t)y * s + x;
            R px = (R)x + uv[2 * pos], py = (R)y + uv[2 * pos + 1];
            tv2[2 * pos] = FN(bil)(tv, 2, 0, w, h, s, px, py);
            tv2[2 * pos + 1] = FN(bil)(tv, 2, 1, w, h, s, px, py);
        }
}

/* warpImage.cu:136-150  warpUV = u * tv2 */
void FN(oracle_k_flowvec)(const R *u, const R *tv2, int w, int h, int s, R *uv)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t pos = (size_t)y * s + x;
            R us = u[pos];
            R tx = tv2[2 * pos], ty = tv2[2 * pos + 1];
            uv[2 * pos] = us * tx;
            uv[2 * pos + 1] = us * ty;
        }
}

/* downscale.cu:10-30 (float) and :98-122 (float2): mean of 4 bilinear taps at
 * +-1/4 coarse pixel around the coarse pixel centre, fetched from the fine image. */
void FN(oracle_k_downscale)(const R *src, int nch, int W, int H, int S,
                            int w, int h, int s, R *out)
{
    R rx = (R)W / (R)w, ry = (R)H / (R)h;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            R cx = ((R)x + (R)0.5) * rx - (R)0.5;
            R cy = ((R)y + (R)0.5) * ry - (R)0.5;
            R ox = (R)0.25 * rx, oy = (R)0.25 * ry;
            for (int c = 0; c < nch; c++) {
                R t = FN(bil)(src, nch, c, W, H, S, cx - ox, cy) +
                      FN(bil)(src, nch, c, W, H, S, cx + ox, cy) +
                      FN(bil)(src, nch, c, W, H, S, cx, cy - oy) +
                      FN(bil)(src, nch, c, W, H, S, cx, cy + oy);
                out[((size_t)y * s + x) * nch + c] = (R)0.25 * t;
            }
        }
}

/* downscale.cu:52-73  point-filtered sample at the coarse pixel centre:
 * texel floor((x+0.5) * W / w), computed exactly in integers. */
void FN(oracle_k_downscale_nn)(const R *src, int W, int H, int S, int w, int h, int s, R *out)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            long fx = ((2L * x + 1) * W) / (2L * w);
            long fy = ((2L * y + 1) * H) / (2L * h);
            out[(size_t)y * s + x] = FN(texel)(src, W, H, S, fx, fy);
        }
}

/* upscale.cu:7-21 (float), :46-61 (float2): bilinear at fine pixel centres of the
 * coarse image, times `scale` (fScale, not the true size ratio). */
void FN(oracle_k_upscale)(const R *src, int nch, int w, int h, int s,
                          int W, int H, int S, R scale, R *out)
{
    R rx = (R)w / (R)W, ry = (R)h / (R)H;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            R px = ((R)x + (R)0.5) * rx - (R)0.5;
            R py = ((R)y + (R)0.5) * ry - (R)0.5;
            for (int c = 0; c < nch; c++)
                out[((size_t)y * S + x) * nch + c] = FN(bil)(src, nch, c, w, h, s, px, py) * scale;
        }
}

/* calcTensor.cu:99-128 (unmasked, mask == NULL) and :245-273 (masked):
 * 3x3 binomial [1 2 1]^T[1 2 1]/16 with mirror edges. */
void FN(oracle_k_gauss)(const R *in, const R *mask, int w, int h, int s, R *out)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t pos = (size_t)y * s + x;
            if (mask && mask[pos] == (R)0) continue;
            R t = (R)0.25 * FN(texel)(in, w, h, s, x, y);
            t += (R)0.0625 * FN(texel)(in, w, h, s, x - 1, y - 1);
            t += (R)0.0625 * FN(texel)(in, w, h, s, x - 1, y + 1);
            t += (R)0.0625 * FN(texel)(in, w, h, s, x + 1, y - 1);
            t += (R)0.0625 * FN(texel)(in, w, h, s, x + 1, y + 1);
            t += (R)0.125 * FN(texel)(in, w, h, s, x - 1, y);
            t += (R)0.125 * FN(texel)(in, w, h, s, x + 1, y);
            t += (R)0.125 * FN(texel)(in, w, h, s, x, y - 1);
            t += (R)0.125 * FN(texel)(in, w, h, s, x, y + 1);
            out[pos] = t;
        }
}

/* calcTensor.cu:7-73 (unmasked) and :154-219 (masked).  The launch covers only
 * columns x < xlimit where xlimit = blockWidth * ceil(h / blockWidth)
 * (calcTensor.cu:81,227 size grid.x by the HEIGHT); other columns keep their value. */
void FN(oracle_k_tensor)(const R *gray, const R *mask, R beta, R gamma,
                         int w, int h, int s, int xlimit, R *at, R *bt, R *ct)
{
    int xe = xlimit < w ? xlimit : w;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < xe; x++) {
            size_t pos = (size_t)y * s + x;
            if (mask && mask[pos] == (R)0) continue;
            /* the first difference of each pair is overwritten (calcTensor.cu:28-39) */
            R gx = FN(texel)(gray, w, h, s, x + 1, y + 1) - FN(texel)(gray, w, h, s, x, y + 1);
            R gy = FN(texel)(gray, w, h, s, x + 1, y + 1) - FN(texel)(gray, w, h, s, x + 1, y);
            R absg = sqrt(gx * gx + gy * gy);
            R nx = gx / absg, ny = gy / absg;
            if (absg < (R)1e-8f) { nx = 1; ny = 0; }
            R tx = ny, ty = -nx;
            R wt = exp(-beta * pow(absg, gamma));
            if (!(wt > (R)1e-8f)) wt = (R)1e-8f;
            at[pos] = wt * nx * nx + tx * tx;
            ct[pos] = wt * nx * ny + tx * ty;
            bt[pos] = wt * ny * ny + ty * ty;
        }
}

/* preconditioning.cu:3-22 (unmasked) and :41-60 (masked) */
void FN(oracle_k_eta)(const R *mask, R alpha0, R alpha1, const R *at, const R *bt, const R *ct,
                      int w, int h, int s, R *etau, R *etav1, R *etav2)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t pos = (size_t)y * s + x;
            if (mask && mask[pos] == (R)0) continue;
            R a = at[pos], b = bt[pos], c = ct[pos];
            etau[pos] = (a * a + b * b + 2 * c * c + (a + c) * (a + c) + (b + c) * (b + c)) * (alpha1 * alpha1);
            etav1[pos] = (alpha1 * alpha1) * (b * b + c * c) + 4 * alpha0 * alpha0;
            etav2[pos] = (alpha1 * alpha1) * (a * a + c * c) + 4 * alpha0 * alpha0;
        }
}

/* updateDualVariables.cu:122-237.  p (2ch), q (4ch) updated in place; grad_v (4ch)
 * written.  Forward differences fall back to backward ones where the right/down
 * neighbour is masked out or off-image. */
void FN(oracle_k_dual_masked)(const R *mask, const R *u_, const R *v_, R alpha0, R alpha1, R sigma,
                              R eta_p, R eta_q, const R *at, const R *bt, const R *ct,
                              int w, int h, int s, R *gradv, R *p, R *q)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            long pos = (long)y * s + x;
            if (mask[pos] == (R)0) continue;
            long right = pos + 1, down = pos + s, left = pos - 1, up = pos - s;
            R mR = (x + 1 >= w) ? (R)0 : mask[right];
            R mD = (y + 1 >= h) ? (R)0 : mask[down];
            R up_ = u_[pos];
            R vx = v_[2 * pos], vy = v_[2 * pos + 1];
            R ux, uy;
            if (mR != (R)0) ux = u_[right] - up_ - vx;
            else ux = up_ - FN(rd)(u_, left) - vx;
            if (mD != (R)0) uy = u_[down] - up_ - vy;
            else uy = up_ - FN(rd)(u_, up) - vy;
            R a = at[pos], b = bt[pos], c = ct[pos];
            R dtx = a * ux + c * uy;
            R dty = c * ux + b * uy;
            R k1 = alpha1 * sigma / eta_p;
            R px = p[2 * pos] + k1 * dtx;
            R py = p[2 * pos + 1] + k1 * dty;
            R rp = sqrt(px * px + py * py);
            if (rp < 1) rp = 1;
            p[2 * pos] = px / rp;
            p[2 * pos + 1] = py / rp;
            R g0, g1, g2, g3;
            if (mR != (R)0) g0 = v_[2 * right] - vx;
            else g0 = vx - (left < 0 ? (R)0 : v_[2 * left]);
            if (mD != (R)0) g1 = v_[2 * down + 1] - vy;
            else g1 = vy - (up < 0 ? (R)0 : v_[2 * up + 1]);
            if (mD != (R)0) g2 = v_[2 * down] - vx;
            else g2 = vx - (up < 0 ? (R)0 : v_[2 * up]);
            if (mR != (R)0) g3 = v_[2 * right + 1] - vy;
            else g3 = vy - (left < 0 ? (R)0 : v_[2 * left + 1]);
            gradv[4 * pos] = g0; gradv[4 * pos + 1] = g1; gradv[4 * pos + 2] = g2; gradv[4 * pos + 3] = g3;
            R k0 = alpha0 * sigma / eta_q;
            R q0 = q[4 * pos] + k0 * g0, q1 = q[4 * pos + 1] + k0 * g1;
            R q2 = q[4 * pos + 2] + k0 * g2, q3 = q[4 * pos + 3] + k0 * g3;
            R rq = sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
            if (rq < 1) rq = 1;
            q[4 * pos] = q0 / rq; q[4 * pos + 1] = q1 / rq; q[4 * pos + 2] = q2 / rq; q[4 * pos + 3] = q3 / rq;
        }
}

/* The unmasked solve's circle gate (updateDualVariables.cu:13-22, thresholdingL1.cu:10-19,
 * updatePrimalVariables.cu:14-23): radius = sqrtf((y - h/2)^2 + (x - w/2)^2) >= w/2.2,
 * in float arithmetic with the correctly rounded square root, as the reference. */
static inline int FN(outside_circle)(int x, int y, int w, int h)
{
    float desired = (float)w / 2.20f;
    float hw = (float)w / 2.0f, hh = (float)h / 2.0f;
    float dy = (float)y - hh, dx = (float)x - hw;
    return sqrtf(dy * dy + dx * dx) >= desired;
}

/* updateDualVariables.cu:4-101 (unmasked; writes p = q = 0 outside the circle) */
void FN(oracle_k_dual)(const R *u_, const R *v_, R alpha0, R alpha1, R sigma, R eta_p, R eta_q,
                       const R *at, const R *bt, const R *ct, int w, int h, int s,
                       R *gradv, R *p, R *q)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            long pos = (long)y * s + x;
            if (FN(outside_circle)(x, y, w, h)) {
                p[2 * pos] = 0; p[2 * pos + 1] = 0;
                q[4 * pos] = 0; q[4 * pos + 1] = 0; q[4 * pos + 2] = 0; q[4 * pos + 3] = 0;
                continue;
            }
            long right = pos + 1, down = pos + s, left = pos - 1, up = pos - s;
            R vx = v_[2 * pos], vy = v_[2 * pos + 1];
            R ux, uy;
            if (x + 1 < w) ux = u_[right] - u_[pos] - vx;
            else ux = u_[pos] - FN(rd)(u_, left) - vx;
            if (y + 1 < h) uy = u_[down] - u_[pos] - vy;
            else uy = u_[pos] - FN(rd)(u_, up) - vy;
            R a = at[pos], b = bt[pos], c = ct[pos];
            R dtx = a * ux + c * uy, dty = c * ux + b * uy;
            R k1 = alpha1 * sigma / eta_p;
            R px = p[2 * pos] + k1 * dtx, py = p[2 * pos + 1] + k1 * dty;
            R rp = sqrt(px * px + py * py);
            if (rp < 1) rp = 1;
            p[2 * pos] = px / rp; p[2 * pos + 1] = py / rp;
            R g0, g1, g2, g3;
            if (x + 1 < w) g0 = v_[2 * right] - vx; else g0 = vx - (left < 0 ? (R)0 : v_[2 * left]);
            if (y + 1 < h) g1 = v_[2 * down + 1] - vy; else g1 = vy - (up < 0 ? (R)0 : v_[2 * up + 1]);
            if (y + 1 < h) g2 = v_[2 * down] - vx; else g2 = vx - (up < 0 ? (R)0 : v_[2 * up]);
            if (x + 1 < w) g3 = v_[2 * right + 1] - vy; else g3 = vy - (left < 0 ? (R)0 : v_[2 * left + 1]);
            gradv[4 * pos] = g0; gradv[4 * pos + 1] = g1; gradv[4 * pos + 2] = g2; gradv[4 * pos + 3] = g3;
            R k0 = alpha0 * sigma / eta_q;
            R q0 = q[4 * pos] + k0 * g0, q1 = q[4 * pos + 1] + k0 * g1;
            R q2 = q[4 * pos + 2] + k0 * g2, q3 = q[4 * pos + 3] + k0 * g3;
            R rq = sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
            if (rq < 1) rq = 1;
            q[4 * pos] = q0 / rq; q[4 * pos + 1] = q1 / rq; q[4 * pos + 2] = q2 / rq; q[4 * pos + 3] = q3 / rq;
        }
}

/* thresholdingL1.cu:234-244 (masked, mask != NULL) and :104-115 (unmasked) */
void FN(oracle_k_tp)(const R *mask, const R *at, const R *bt, const R *ct, const R *p,
                     int w, int h, int s, R *Tp)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t pos = (size_t)y * s + x;
            if (mask && mask[pos] == (R)0) continue;
            R a = at[pos], b = bt[pos], c = ct[pos], px = p[2 * pos], py = p[2 * pos + 1];
            Tp[2 * pos] = a * px + c * py;
            Tp[2 * pos + 1] = c * px + b * py;
        }
}

/* L1 proximal step shared by both thresholding kernels: thresholdingL1.cu:180-215
 * (masked) / :47-83 (unmasked).  Linearised at the current u (not u_last). */
static inline R FN(thresh_core)(R divp, R tau, R etau, R u_pos, R upos, R Iu, R Iz, R lambda)
{
    R te = (etau == (R)0) ? tau : tau / etau;
    R uhat = u_pos + te * divp;
    R dun = uhat - upos;
    R rho = Iu * dun + Iz;
    R upper = lambda * te * (Iu * Iu);
    R lower = -lambda * te * (Iu * Iu);
    R du = 0; /* the reference leaves du uninitialised when rho is NaN */
    if ((rho <= upper) && (rho >= lower)) {
        if (Iu == (R)0) du = dun;
        else du = dun - rho / Iu;
    } else if (rho < lower) {
        du = dun + lambda * te * Iu;
    } else if (rho > upper) {
        du = dun - lambda * te * Iu;
    }
    return upos + du;
}

/* thresholdingL1.cu:128-216 */
void FN(oracle_k_thresh_masked)(const R *Tp, const R *u_, const R *Iu, const R *Iz, const R *mask,
                                R lambda, R tau, const R *etau, const R *u, R *us,
                                int w, int h, int s)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            long pos = (long)y * s + x;
            if (mask[pos] == (R)0) continue;
            long right = pos + 1, down = pos + s, left = pos - 1, up = pos - s;
            R mR = (x + 1 >= w) ? (R)0 : mask[right];
            R mL = (x - 1 < 0) ? (R)0 : mask[left];
            R mD = (y + 1 >= h) ? (R)0 : mask[down];
            R mU = (y - 1 < 0) ? (R)0 : mask[up];
            R dx, dy;
            if (mL != (R)0 && mR != (R)0) dx = Tp[2 * pos] - Tp[2 * left];
            else if (mR == (R)0) dx = -(left < 0 ? (R)0 : Tp[2 * left]);
            else dx = Tp[2 * pos];
            if (mU != (R)0 && mD != (R)0) dy = Tp[2 * pos + 1] - Tp[2 * up + 1];
            else if (mD == (R)0) dy = -(up < 0 ? (R)0 : Tp[2 * up + 1]);
            else dy = Tp[2 * pos + 1];
            us[pos] = FN(thresh_core)(dx + dy, tau, etau[pos], u_[pos], u[pos], Iu[pos], Iz[pos], lambda);
        }
}

/* thresholdingL1.cu:3-86 (unmasked; us = 0 outside the circle) */
void FN(oracle_k_thresh)(const R *Tp, const R *u_, const R *Iu, const R *Iz,
                         R lambda, R tau, const R *etau, const R *u, R *us, int w, int h, int s)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            long pos = (long)y * s + x;
            if (FN(outside_circle)(x, y, w, h)) { us[pos] = 0; continue; }
            long left = pos - 1, up = pos - s;
            R dx, dy;
            if (x - 1 >= 0) dx = Tp[2 * pos] - Tp[2 * left];
            else if (x == w - 1) dx = -(left < 0 ? (R)0 : Tp[2 * left]);
            else dx = Tp[2 * pos];
            if (y - 1 >= 0) dy = Tp[2 * pos + 1] - Tp[2 * up + 1];
            else if (y == h - 1) dy = -(up < 0 ? (R)0 : Tp[2 * up + 1]);
            else dy = Tp[2 * pos + 1];
            us[pos] = FN(thresh_core)(dx + dy, tau, etau[pos], u_[pos], u[pos], Iu[pos], Iz[pos], lambda);
        }
}

/* v step + over-relaxation shared by both primal kernels: updatePrimalVariables.cu:211-244
 * (masked) / :76-110 (unmasked).  The eta_v2 == 0 branch writes .x (as the reference). */
static inline void FN(primal_core)(long pos, R qcx, R qcy, R qcz, R qcw, R qwx, R qww, R qny, R qnz,
                                   const R *p, const R *at, const R *bt, const R *ct,
                                   R tau, const R *etav1, const R *etav2, R alpha0, R alpha1, R mu,
                                   const R *u_, const R *v_, const R *u, R *v, R *u_s, R *v_s)
{
    R dqx = (qcx - qwx) + (qcz - qnz);
    R dqy = (qcw - qww) + (qcy - qny);
    R px = p[2 * pos], py = p[2 * pos + 1];
    R a = at[pos], b = bt[pos], c = ct[pos];
    R tx = a * px + c * py, ty = c * px + b * py;
    R e1 = etav1[pos], e2 = etav2[pos];
    R v_x = v_[2 * pos], v_y = v_[2 * pos + 1];
    R vx, vy;
    if (e1 != (R)0) vx = v_x + (tau / e1) * (alpha1 * tx + alpha0 * dqx);
    else vx = v_x + tau * (alpha1 * tx + alpha0 * dqx);
    if (e2 != (R)0) vy = v_y + (tau / e2) * (alpha1 * ty + alpha0 * dqy);
    else {
        /* the reference leaves vpos.y uninitialised here; unreachable for alpha0 != 0 */
        vx = v_x + tau * (alpha1 * tx + alpha0 * dqx);
        vy = 0;
    }
    v[2 * pos] = vx; v[2 * pos + 1] = vy;
    u_s[pos] = u[pos] + mu * (u[pos] - u_[pos]);
    v_s[2 * pos] = vx + mu * (vx - v_x);
    v_s[2 * pos + 1] = vy + mu * (vy - v_y);
}

/* updatePrimalVariables.cu:138-245 */
void FN(oracle_k_primal_masked)(const R *mask, const R *u_, const R *v_, const R *p, const R *q,
                                const R *at, const R *bt, const R *ct, R tau,
                                const R *etav1, const R *etav2, R alpha0, R alpha1, R mu,
                                const R *u, R *v, R *u_s, R *v_s, int w, int h, int s)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            long pos = (long)y * s + x;
            if (mask[pos] == (R)0) continue;
            long right = pos + 1, down = pos + s, left = pos - 1, up = pos - s;
            R mR = (x + 1 >= w) ? (R)0 : mask[right];
            R mL = (x - 1 < 0) ? (R)0 : mask[left];
            R mD = (y + 1 >= h) ? (R)0 : mask[down];
            R mU = (y - 1 < 0) ? (R)0 : mask[up];
            R qcx = 0, qcw = 0, qcy = 0, qcz = 0;
            if (mR != (R)0) { qcx = q[4 * pos]; qcw = q[4 * pos + 3]; }
            if (mD != (R)0) { qcy = q[4 * pos + 1]; qcz = q[4 * pos + 2]; }
            R qwx = 0, qww = 0, qny = 0, qnz = 0;
            if (mL != (R)0) { qwx = q[4 * left]; qww = q[4 * left + 3]; }
            if (mU != (R)0) { qny = q[4 * up + 1]; qnz = q[4 * up + 2]; }
            FN(primal_core)(pos, qcx, qcy, qcz, qcw, qwx, qww, qny, qnz, p, at, bt, ct, tau,
                            etav1, etav2, alpha0, alpha1, mu, u_, v_, u, v, u_s, v_s);
        }
}

/* updatePrimalVariables.cu:3-113 (unmasked; u_s = v_s = 0 outside the circle) */
void FN(oracle_k_primal)(const R *u_, const R *v_, const R *p, const R *q,
                         const R *at, const R *bt, const R *ct, R tau,
                         const R *etav1, const R *etav2, R alpha0, R alpha1, R mu,
                         const R *u, R *v, R *u_s, R *v_s, int w, int h, int s)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            long pos = (long)y * s + x;
            if (FN(outside_circle)(x, y, w, h)) {
                u_s[pos] = 0; v_s[2 * pos] = 0; v_s[2 * pos + 1] = 0;
                continue;
            }
            long left = pos - 1, up = pos - s;
            R qcx = 0, qcw = 0, qcy = 0, qcz = 0;
            if (x != w - 1) { qcx = q[4 * pos]; qcw = q[4 * pos + 3]; }
            if (y != h - 1) { qcy = q[4 * pos + 1]; qcz = q[4 * pos + 2]; }
            R qwx = 0, qww = 0, qny = 0, qnz = 0;
            if (x - 1 >= 0) { qwx = q[4 * left]; qww = q[4 * left + 3]; }
            if (y - 1 >= 0) { qny = q[4 * up + 1]; qnz = q[4 * up + 2]; }
            FN(primal_core)(pos, qcx, qcy, qcz, qcw, qwx, qww, qny, qnz, p, at, bt, ct, tau,
                            etav1, etav2, alpha0, alpha1, mu, u_, v_, u, v, u_s, v_s);
        }
}

/* medianFilter.cu:264-319: 13th smallest of the 5x5 window.  Exactly one of
 * row/col out of range -> edge replicate; both out of range -> 0. */
void FN(oracle_k_median5)(const R *u, int w, int h, int s, R *out)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            R m[25];
            for (int j = 0; j < 5; j++)
                for (int i = 0; i < 5; i++) {
                    int col = x + i - 2, row = y + j - 2;
                    int cin = col >= 0 && col < w, rin = row >= 0 && row < h;
                    R val = 0;
                    if (cin && rin) val = u[(size_t)row * s + col];
                    else if (col < 0 && rin) val = u[(size_t)row * s];
                    else if (col >= w && rin) val = u[(size_t)row * s + w - 1];
                    else if (cin && row < 0) val = u[col];
                    else if (cin && row >= h) val = u[(size_t)(h - 1) * s + col];
                    m[j * 5 + i] = val;
                }
            /* partial selection sort, as medianFilter.cu:307-316 */
            for (int j = 0; j < 13; j++)
                for (int i = j + 1; i < 25; i++)
                    if (m[j] > m[i]) { R t = m[j]; m[j] = m[i]; m[i] = t; }
            out[(size_t)y * s + x] = m[12];
        }
}

/* scalar.cu:129-141, limitRange.cu:3-24, scalar.cu:77-89, copyClone.cu:3-30, scalar.cu:102-115 */
void FN(oracle_k_sub)(const R *a, const R *b, int w, int h, int s, R *d)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) { size_t p = (size_t)y * s + x; d[p] = a[p] - b[p]; }
}
void FN(oracle_k_limit)(const R *src, R lim, int w, int h, int s, R *dst)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t p = (size_t)y * s + x;
            dst[p] = (src[p] > lim) ? lim : src[p];
        }
}
void FN(oracle_k_add)(const R *a, const R *b, int nch, int w, int h, int s, R *d)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w * nch; x++) {
            size_t p = (size_t)y * s * nch + x;
            d[p] = a[p] + b[p];
        }
}
void FN(oracle_k_clone)(R *dst, int nch, int w, int h, int s, const R *src)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w * nch; x++) {
            size_t p = (size_t)y * s * nch + x;
            dst[p] = src[p];
        }
}

/* conversion.cu:3-69: Kannala-Brandt back-projection of both frames (5 Newton
 * steps from theta = 0), Z from the x components, X = (x0 Z, y0 Z, Z), depth = |X|.
 * (u - focalx)/cx: the reference passes the principal point as `focal` and vice versa. */
void FN(oracle_k_convert_kb)(const R *uv, R focalx, R focaly, R cx, R cy,
                             R d1, R d2, R d3, R d4, R t1, R t2, R t3,
                             R *X, R *depth, int w, int h, int s)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t pos = (size_t)y * s + x;
            R u0 = (R)x, v0 = (R)y;
            R xp0 = (u0 - focalx) / cx, yp0 = (v0 - focaly) / cy;
            R u1 = u0 + uv[2 * pos], v1 = v0 + uv[2 * pos + 1];
            R xp1 = (u1 - focalx) / cx, yp1 = (v1 - focaly) / cy;
            R r0 = sqrt(xp0 * xp0 + yp0 * yp0);
            R th = 0;
            for (int it = 0; it < 5; it++) {
                R t2_ = th * th;
                R td = th + d1 * pow(th, 3) + d2 * pow(th, 5) + d3 * pow(th, 7) + d4 * pow(th, 9);
                R dtd = 1 + 3 * d1 * t2_ + 5 * d2 * pow(th, 4) + 7 * d3 * pow(th, 6) + 9 * d4 * pow(th, 8);
                th = th - (r0 - td) / (-dtd);
            }
            R x0o = tan(th) * xp0 / r0, y0o = tan(th) * yp0 / r0;
            R r1 = sqrt(xp1 * xp1 + yp1 * yp1);
            th = 0;
            for (int it = 0; it < 5; it++) {
                R t2_ = th * th;
                R td = th + d1 * pow(th, 3) + d2 * pow(th, 5) + d3 * pow(th, 7) + d4 * pow(th, 9);
                R dtd = 1 + 3 * d1 * t2_ + 5 * d2 * pow(th, 4) + 7 * d3 * pow(th, 6) + 9 * d4 * pow(th, 8);
                th = th - (r1 - td) / (-dtd);
            }
            R x1o = tan(th) * xp1 / r1;
            R Z = (t1 - x1o * t3) / (x1o - x0o);
            (void)t2;
            R X0 = x0o * Z, X1 = y0o * Z, X2 = Z;
            X[3 * pos] = X0; X[3 * pos + 1] = X1; X[3 * pos + 2] = X2;
            depth[pos] = sqrt(X0 * X0 + X1 * X1 + X2 * X2);
        }
}

/* flowToHsv.cu:3-111 (ComputeColorKernel): flow colour coding.  (du, dv) = uv / scale;
 * angle a = atan2(-dv, -du) / 3.14159 picks a position fk = (a + 1) / 2 * 54 + 1 on the
 * 55-entry Middlebury colour wheel (segments red-yellow 15, yellow-green 6, green-cyan
 * 4, cyan-blue 11, blue-magenta 13, magenta-red 6; entry i of a rising segment of length
 * n is floor(255 i / n)), interpolated between floor(fk) and floor(fk) + 1 (wrapping
 * 56 -> 1); the reference indexes its 55-entry table with these 1-based positions, so
 * position 55 (angle within 2.7e-6 rad of +pi) reads past its table and is taken as
 * entry 0 here.  Magnitude <= 1 blends towards white (1 - r (1 - c)), larger ones are
 * darkened (0.75 c).  Output per pixel: (B, G, R) in [0, 1], the CV_32FC3 layout of the
 * reference's host Mat. */
static void FN(wheel_rgb)(int k, R *r, R *g, R *b)
{
    static const int seg[6] = {15, 6, 4, 11, 13, 6};
    k = ((k % 55) + 55) % 55;
    int start = 0, si = 0;
    while (k >= start + seg[si]) start += seg[si++];
    const int i = k - start, n = seg[si];
    const int up = (255 * i) / n, down = 255 - (255 * i) / n;
    int rr = 0, gg = 0, bb = 0;
    switch (si) {
    case 0: rr = 255; gg = up; break;            /* red -> yellow */
    case 1: rr = down; gg = 255; break;          /* yellow -> green */
    case 2: gg = 255; bb = up; break;            /* green -> cyan */
    case 3: gg = down; bb = 255; break;          /* cyan -> blue */
    case 4: rr = up; bb = 255; break;            /* blue -> magenta */
    default: rr = 255; bb = down; break;         /* magenta -> red */
    }
    *r = (R)rr; *g = (R)gg; *b = (R)bb;
}

void FN(oracle_wheel)(int k, R *rgb) { FN(wheel_rgb)(k, &rgb[0], &rgb[1], &rgb[2]); }

void FN(oracle_k_flow_hsv)(const R *uv, R scale, R *bgr, int w, int h, int s)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t pos = (size_t)y * s + x;
            R du = uv[2 * pos] / scale, dv = uv[2 * pos + 1] / scale;
            R rad = sqrt(du * du + dv * dv);
            R a = atan2(-dv, -du) / (R)3.14159;
            R fk = (a + 1) / 2 * (R)54 + 1;
            int k0 = (int)floor(fk);
            int k1 = k0 + 1;
            if (k1 == 56) k1 = 1;
            R f = fk - (R)k0;
            R r0, g0, b0, r1, g1, b1;
            FN(wheel_rgb)(k0, &r0, &g0, &b0);
            FN(wheel_rgb)(k1, &r1, &g1, &b1);
            R cr = (1 - f) * (r0 / 255) + f * (r1 / 255);
            R cg = (1 - f) * (g0 / 255) + f * (g1 / 255);
            R cb = (1 - f) * (b0 / 255) + f * (b1 / 255);
            if (rad <= 1) {
                cr = 1 - rad * (1 - cr);
                cg = 1 - rad * (1 - cg);
                cb = 1 - rad * (1 - cb);
            } else {
                cr = cr * (R)0.75;
                cg = cg * (R)0.75;
                cb = cb * (R)0.75;
            }
            bgr[3 * pos] = cb; bgr[3 * pos + 1] = cg; bgr[3 * pos + 2] = cr;
        }
}

/* computeDerivatives.cu:88-129: 5-point derivative along the normalised level
 * trajectory t = tv/|tv| (|tv| = 0 gives NaN, as the reference), averaged over
 * I0 and the warped I1; Iz = I1w - I0 at the texel centre. */
void FN(oracle_deriv)(const R *I0, const R *I1w, const R *tv, int w, int h, int s, R *Iu, R *Iz)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            size_t pos = (size_t)y * s + x;
            R vx = tv[2 * pos], vy = tv[2 * pos + 1];
            R r = sqrt(vx * vx + vy * vy);
            R dx = vx / r, dy = vy / r;
            R X = (R)x, Y = (R)y;
            R t0 = FN(bil)(I0, 1, 0, w, h, s, X - 2 * dx, Y - 2 * dy);
            t0 -= FN(bil)(I0, 1, 0, w, h, s, X - dx, Y - dy) * 8;
            t0 += FN(bil)(I0, 1, 0, w, h, s, X + dx, Y + dy) * 8;
            t0 -= FN(bil)(I0, 1, 0, w, h, s, X + 2 * dx, Y + 2 * dy);
            t0 /= 12;
            R t1 = FN(bil)(I1w, 1, 0, w, h, s, X - 2 * dx, Y - 2 * dy);
            t1 -= FN(bil)(I1w, 1, 0, w, h, s, X - dx, Y - dy) * 8;
            t1 += FN(bil)(I1w, 1, 0, w, h, s, X + dx, Y + dy) * 8;
            t1 -= FN(bil)(I1w, 1, 0, w, h, s, X + 2 * dx, Y + 2 * dy);
            t1 /= 12;
            Iu[pos] = (t0 + t1) * (R)0.5;
            Iz[pos] = I1w[pos] - I0[pos];
        }
}

/* Energy monitor of a warp (no reference counterpart; the GPU solver's FP64 energy and
 * residual records): over the processed pixels (mask != 0, or inside the circle when
 * mask == NULL), with the dual step's gradients (updateDualVariables.cu:152-212: forward
 * differences, backward where the right / lower neighbour is outside the mask or the
 * image, flat index < 0 reads 0):
 *   term = lambda |Iu (u - u_last) + Iz| + alpha1 |T (grad u - v)| + alpha0 |grad v|,
 *   T = [[a, c], [c, b]];  residual = |u - u_prev|.
 * out = {sum of terms, sum of data terms, sum of residuals, pixels with a non-finite term
 * or residual (left out of the sums), processed pixels}.  Evaluated in double in row
 * order. */
void FN(oracle_k_energy)(const R *mask, const R *u, const R *uprev, const R *v, const R *Iu, const R *Iz,
                         const R *ulast, const R *at, const R *bt, const R *ct, int w, int h, int s,
                         R lambda, R alpha0, R alpha1, double *out)
{
    double se = 0, sd = 0, sr = 0, nf = 0, np = 0;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            long pos = (long)y * s + x;
            int hasR, hasD;
            if (mask) {
                if (mask[pos] == (R)0) continue;
                hasR = x + 1 < w && mask[pos + 1] != (R)0;
                hasD = y + 1 < h && mask[pos + s] != (R)0;
            } else {
                if (FN(outside_circle)(x, y, w, h)) continue;
                hasR = x + 1 < w;
                hasD = y + 1 < h;
            }
            double u0 = (double)u[pos], v0x = (double)v[2 * pos], v0y = (double)v[2 * pos + 1];
            double ux, uy, g0, g1, g2, g3;
            if (hasR) {
                ux = ((double)u[pos + 1] - u0) - v0x;
                g0 = (double)v[2 * (pos + 1)] - v0x;
                g3 = (double)v[2 * (pos + 1) + 1] - v0y;
            } else {
                double uL = pos >= 1 ? (double)u[pos - 1] : 0.0;
                double vLx = pos >= 1 ? (double)v[2 * (pos - 1)] : 0.0, vLy = pos >= 1 ? (double)v[2 * (pos - 1) + 1] : 0.0;
                ux = (u0 - uL) - v0x;
                g0 = v0x - vLx;
                g3 = v0y - vLy;
            }
            if (hasD) {
                uy = ((double)u[pos + s] - u0) - v0y;
                g1 = (double)v[2 * (pos + s) + 1] - v0y;
                g2 = (double)v[2 * (pos + s)] - v0x;
            } else {
                double uU = pos >= s ? (double)u[pos - s] : 0.0;
                double vUx = pos >= s ? (double)v[2 * (pos - s)] : 0.0, vUy = pos >= s ? (double)v[2 * (pos - s) + 1] : 0.0;
                uy = (u0 - uU) - v0y;
                g1 = v0y - vUy;
                g2 = v0x - vUx;
            }
            double a = (double)at[pos], b = (double)bt[pos], c = (double)ct[pos];
            double tx = a * ux + c * uy, ty = c * ux + b * uy;
            double rho = (double)Iu[pos] * (u0 - (double)ulast[pos]) + (double)Iz[pos];
            double data = (double)lambda * fabs(rho);
            double term = data + (double)alpha1 * sqrt(tx * tx + ty * ty) +
                          (double)alpha0 * sqrt(g0 * g0 + g1 * g1 + g2 * g2 + g3 * g3);
            double res = fabs(u0 - (double)uprev[pos]);
            np += 1;
            if (isfinite(term) && isfinite(res)) { se += term; sd += data; sr += res; }
            else nf += 1;
        }
    out[0] = se; out[1] = sd; out[2] = sr; out[3] = nf; out[4] = np;
}

#ifndef ORACLE_MON_WARPS
#define ORACLE_MON_WARPS 128
#endif

/* ------------------------------------------------------------ solver state */

typedef struct FN(oracle_s) {
    int width, height, stride, nLevels, nWarpIters, nSolverIters;
    int blockWidth, strideAlign;
    R beta, gamma, alpha0, alpha1, timestep_lambda, lambda, fScale, limitRange;
    int pW[ORACLE_MAXL], pH[ORACLE_MAXL], pS[ORACLE_MAXL];
    R *pI0[ORACLE_MAXL], *pI1[ORACLE_MAXL], *pTv[ORACLE_MAXL], *pMask[ORACLE_MAXL];
    /* level-0-sized work buffers (stereotgv.cu:106-161) */
    R *i1cal, *i1warp, *Iu, *Iz, *u, *du, *us, *warpUV, *dwarpUV, *warpUVs, *tv2, *cv;
    R *a, *b, *c, *etau, *etav1, *etav2, *p, *q, *u_, *u_s, *u_last, *v, *v_, *v_s, *gradv, *Tp, *i0smooth;
    R *X, *depth;
    double mon[ORACLE_MAXL][ORACLE_MON_WARPS][5]; /* oracle_k_energy of every (level, warp) */
    R *work[40];
    size_t worksz[40];
    int nwork;
} FN(oracle_t);

static int FN(align_up)(int n, int m) { int r = n % m; return r ? n + m - r : n; }

static R *FN(walloc)(FN(oracle_t) *o, size_t n)
{
    R *p = (R *)calloc(n, sizeof(R));
    o->work[o->nwork] = p;
    o->worksz[o->nwork] = n;
    o->nwork++;
    return p;
}

/* stereotgv.cu:44-168 (initialize), with the constructor's block width and
 * stride alignment (stereotgv.cu:32-42). */
FN(oracle_t) *FN(oracle_create)(int width, int height, float beta, float gamma, float alpha0,
                                float alpha1, float timestep_lambda, float lambda, int nLevels,
                                float fScale, int nWarpIters, int nSolverIters,
                                int blockWidth, int strideAlign)
{
    if (nLevels < 1 || nLevels > ORACLE_MAXL || width < 1 || height < 1) return NULL;
    FN(oracle_t) *o = (FN(oracle_t) *)calloc(1, sizeof(FN(oracle_t)));
    o->width = width; o->height = height;
    o->blockWidth = blockWidth > 0 ? blockWidth : 32;
    o->strideAlign = strideAlign > 0 ? strideAlign : 32;
    o->stride = FN(align_up)(width, o->strideAlign);
    o->beta = beta; o->gamma = gamma; o->alpha0 = alpha0; o->alpha1 = alpha1;
    o->timestep_lambda = timestep_lambda; o->lambda = lambda; o->fScale = fScale;
    o->nLevels = nLevels; o->nWarpIters = nWarpIters; o->nSolverIters = nSolverIters;
    o->limitRange = (R)1.0f;
    int nh = height, nw = width, ns = FN(align_up)(width, o->strideAlign);
    for (int l = 0; l < nLevels; l++) {
        o->pW[l] = nw; o->pH[l] = nh; o->pS[l] = ns;
        size_t n = (size_t)ns * nh;
        o->pI0[l] = (R *)calloc(n, sizeof(R));
        o->pI1[l] = (R *)calloc(n, sizeof(R));
        o->pTv[l] = (R *)calloc(2 * n, sizeof(R));
        o->pMask[l] = (R *)calloc(n, sizeof(R));
        nh = (int)((float)nh / fScale);
        nw = (int)((float)nw / fScale);
        ns = FN(align_up)(nw, o->strideAlign);
        if (l + 1 < nLevels && (nw < 1 || nh < 1)) return NULL;
    }
    size_t N = (size_t)o->stride * height;
    o->i1cal = (R *)calloc(N, sizeof(R)); o->i1warp = FN(walloc)(o, N);
    o->Iu = FN(walloc)(o, N); o->Iz = FN(walloc)(o, N);
    o->u = FN(walloc)(o, N); o->du = FN(walloc)(o, N); o->us = FN(walloc)(o, N);
    o->warpUV = FN(walloc)(o, 2 * N); o->dwarpUV = FN(walloc)(o, 2 * N); o->warpUVs = FN(walloc)(o, 2 * N);
    o->tv2 = FN(walloc)(o, 2 * N);
    o->a = FN(walloc)(o, N); o->b = FN(walloc)(o, N); o->c = FN(walloc)(o, N);
    o->etau = FN(walloc)(o, N); o->etav1 = FN(walloc)(o, N); o->etav2 = FN(walloc)(o, N);
    o->p = FN(walloc)(o, 2 * N); o->q = FN(walloc)(o, 4 * N);
    o->u_ = FN(walloc)(o, N); o->u_s = FN(walloc)(o, N); o->u_last = FN(walloc)(o, N);
    o->v = FN(walloc)(o, 2 * N); o->v_ = FN(walloc)(o, 2 * N); o->v_s = FN(walloc)(o, 2 * N);
    o->gradv = FN(walloc)(o, 4 * N); o->Tp = FN(walloc)(o, 2 * N); o->i0smooth = FN(walloc)(o, N);
    o->cv = (R *)calloc(2 * N, sizeof(R));
    o->X = (R *)calloc(3 * N, sizeof(R));
    o->depth = (R *)calloc(N, sizeof(R));
    return o;
}

void FN(oracle_destroy)(FN(oracle_t) *o)
{
    if (!o) return;
    for (int l = 0; l < o->nLevels; l++) { free(o->pI0[l]); free(o->pI1[l]); free(o->pTv[l]); free(o->pMask[l]); }
    for (int i = 0; i < o->nwork; i++) free(o->work[i]);
    free(o->i1cal);
    free(o->cv); free(o->X); free(o->depth);
    free(o);
}

void FN(oracle_set_limit_range)(FN(oracle_t) *o, float lim) { o->limitRange = (R)lim; }

void FN(oracle_level_dims)(const FN(oracle_t) *o, int level, int *w, int *h, int *s)
{
    *w = o->pW[level]; *h = o->pH[level]; *s = o->pS[level];
}

/* stereotgv.cu:241-252: zero right-padding to the stride, then nearest-neighbour pyramid */
void FN(oracle_copy_mask)(FN(oracle_t) *o, const float *mask, int pitch)
{
    R *m0 = o->pMask[0];
    for (int y = 0; y < o->height; y++)
        for (int x = 0; x < o->stride; x++)
            m0[(size_t)y * o->stride + x] = x < o->width ? (R)mask[(size_t)y * pitch + x] : (R)0;
    for (int l = 1; l < o->nLevels; l++)
        FN(oracle_k_downscale_nn)(o->pMask[l - 1], o->pW[l - 1], o->pH[l - 1], o->pS[l - 1],
                                  o->pW[l], o->pH[l], o->pS[l], o->pMask[l]);
}

/* stereotgv.cu:171-199: trajectory field pyramid by 4-tap box downscale (no
 * renormalisation), calibration field kept at level 0.  The backward field is
 * computed by the reference but never read, so it is omitted here. */
void FN(oracle_load_vector_fields)(FN(oracle_t) *o, const float *tv, int tv_pitch,
                                   const float *cv, int cv_pitch)
{
    R *t0 = o->pTv[0];
    for (int y = 0; y < o->height; y++)
        for (int x = 0; x < o->stride; x++) {
            size_t d = ((size_t)y * o->stride + x) * 2;
            int in = x < o->width;
            t0[d] = in ? (R)tv[(size_t)y * tv_pitch + 2 * x] : (R)0;
            t0[d + 1] = in ? (R)tv[(size_t)y * tv_pitch + 2 * x + 1] : (R)0;
            o->cv[d] = in ? (R)cv[(size_t)y * cv_pitch + 2 * x] : (R)0;
            o->cv[d + 1] = in ? (R)cv[(size_t)y * cv_pitch + 2 * x + 1] : (R)0;
        }
    for (int l = 1; l < o->nLevels; l++)
        FN(oracle_k_downscale)(o->pTv[l - 1], 2, o->pW[l - 1], o->pH[l - 1], o->pS[l - 1],
                               o->pW[l], o->pH[l], o->pS[l], o->pTv[l]);
}

/* stereotgv.cu:227-230 (CV_32F path): copied as-is, right-padded with 0 */
void FN(oracle_copy_images_f32)(FN(oracle_t) *o, const float *i0, const float *i1, int pitch)
{
    for (int y = 0; y < o->height; y++)
        for (int x = 0; x < o->stride; x++) {
            size_t d = (size_t)y * o->stride + x;
            o->pI0[0][d] = x < o->width ? (R)i0[(size_t)y * pitch + x] : (R)0;
            o->pI1[0][d] = x < o->width ? (R)i1[(size_t)y * pitch + x] : (R)0;
        }
}

/* stereotgv.cu:201-213, 220-226 (CV_8U path) */
void FN(oracle_copy_images_u8)(FN(oracle_t) *o, const uint8_t *i0, const uint8_t *i1, int pitch)
{
    FN(oracle_k_cv8u)(i0, pitch, o->pI0[0], o->width, o->height, o->stride);
    FN(oracle_k_cv8u)(i1, pitch, o->pI1[0], o->width, o->height, o->stride);
}

/* stereotgv.cu:231-237 (CV_8UC3 path); pitch in bytes */
void FN(oracle_copy_images_u8c3)(FN(oracle_t) *o, const uint8_t *i0, const uint8_t *i1, int pitch)
{
    FN(oracle_k_cv8uc3)(i0, pitch, o->pI0[0], o->width, o->height, o->stride);
    FN(oracle_k_cv8uc3)(i1, pitch, o->pI1[0], o->width, o->height, o->stride);
}

#define SWAP_PTR(a, b) do { R *t_ = (a); (a) = (b); (b) = t_; } while (0)

static void FN(zero_work)(FN(oracle_t) *o)
{
    for (int i = 0; i < o->nwork; i++) memset(o->work[i], 0, o->worksz[i] * sizeof(R));
    /* the calibration target, whichever buffer currently plays that role */
    memset(o->i1cal, 0, (size_t)o->stride * o->height * sizeof(R));
    for (int l = 1; l < o->nLevels; l++) {
        size_t n = (size_t)o->pS[l] * o->pH[l];
        memset(o->pI0[l], 0, n * sizeof(R));
        memset(o->pI1[l], 0, n * sizeof(R));
    }
}

/* One level of the solve, from the tensor to the end of the last warp
 * (stereotgv.cu:409-494 masked, :275-364 unmasked). */
static void FN(solve_level)(FN(oracle_t) *o, int masked, int l, int warps, int iters)
{
    int w = o->pW[l], h = o->pH[l], s = o->pS[l];
    const R *mask = o->pMask[l];
    size_t N = (size_t)o->stride * o->height;
    memset(o->a, 0, N * sizeof(R)); memset(o->b, 0, N * sizeof(R)); memset(o->c, 0, N * sizeof(R));
    memset(o->etau, 0, N * sizeof(R)); memset(o->etav1, 0, N * sizeof(R)); memset(o->etav2, 0, N * sizeof(R));
    memset(o->i0smooth, 0, N * sizeof(R));
    R eta_p = 3, eta_q = 2;
    int xlimit = o->blockWidth * ((h + o->blockWidth - 1) / o->blockWidth);
    if (masked) {
        FN(oracle_k_gauss)(o->pI0[l], mask, w, h, s, o->i0smooth);
        FN(oracle_k_tensor)(o->i0smooth, mask, o->beta, o->gamma, w, h, s, xlimit, o->a, o->b, o->c);
        FN(oracle_k_eta)(mask, o->alpha0, o->alpha1, o->a, o->b, o->c, w, h, s, o->etau, o->etav1, o->etav2);
    } else {
        FN(oracle_k_gauss)(o->pI0[l], NULL, w, h, s, o->i0smooth);
        FN(oracle_k_tensor)(o->i0smooth, NULL, o->beta, o->gamma, w, h, s, xlimit, o->a, o->b, o->c);
        FN(oracle_k_eta)(NULL, o->alpha0, o->alpha1, o->a, o->b, o->c, w, h, s, o->etau, o->etav1, o->etav2);
    }
    for (int wi = 0; wi < warps; wi++) {
        memset(o->p, 0, 2 * N * sizeof(R)); memset(o->q, 0, 4 * N * sizeof(R));
        memset(o->v, 0, 2 * N * sizeof(R));
        FN(oracle_k_clone)(o->v_, 2, w, h, s, o->v);
        memset(o->gradv, 0, 4 * N * sizeof(R)); memset(o->du, 0, N * sizeof(R));
        FN(oracle_k_find_tv2)(o->warpUV, o->pTv[l], w, h, s, o->tv2);
        FN(oracle_k_warp)(o->pI1[l], w, h, s, o->warpUV, o->i1warp);
        FN(oracle_deriv)(o->pI0[l], o->i1warp, o->pTv[l], w, h, s, o->Iu, o->Iz);
        FN(oracle_k_clone)(o->u_last, 1, w, h, s, o->u);
        /* host step schedule in float (:449-477) */
        float tau = 1.0f, sigma = 1.0f / tau;
        for (int it = 0; it < iters; it++) {
            float mu;
            if (sigma < 1000.0f) mu = 1.0f / sqrtf(1 + 0.7f * tau * (float)o->timestep_lambda);
            else mu = 1;
            if (masked) {
                FN(oracle_k_dual_masked)(mask, o->u_, o->v_, o->alpha0, o->alpha1, (R)sigma, eta_p, eta_q,
                                         o->a, o->b, o->c, w, h, s, o->gradv, o->p, o->q);
                FN(oracle_k_tp)(mask, o->a, o->b, o->c, o->p, w, h, s, o->Tp);
                FN(oracle_k_thresh_masked)(o->Tp, o->u_, o->Iu, o->Iz, mask, o->lambda, (R)tau, o->etau,
                                           o->u, o->us, w, h, s);
                SWAP_PTR(o->u, o->us);
                FN(oracle_k_primal_masked)(mask, o->u_, o->v_, o->p, o->q, o->a, o->b, o->c, (R)tau,
                                           o->etav1, o->etav2, o->alpha0, o->alpha1, (R)mu, o->u,
                                           o->v, o->u_s, o->v_s, w, h, s);
            } else {
                FN(oracle_k_dual)(o->u_, o->v_, o->alpha0, o->alpha1, (R)sigma, eta_p, eta_q,
                                  o->a, o->b, o->c, w, h, s, o->gradv, o->p, o->q);
                FN(oracle_k_tp)(NULL, o->a, o->b, o->c, o->p, w, h, s, o->Tp);
                FN(oracle_k_thresh)(o->Tp, o->u_, o->Iu, o->Iz, o->lambda, (R)tau, o->etau,
                                    o->u, o->us, w, h, s);
                SWAP_PTR(o->u, o->us);
                FN(oracle_k_primal)(o->u_, o->v_, o->p, o->q, o->a, o->b, o->c, (R)tau,
                                    o->etav1, o->etav2, o->alpha0, o->alpha1, (R)mu, o->u,
                                    o->v, o->u_s, o->v_s, w, h, s);
            }
            SWAP_PTR(o->u_, o->u_s);
            SWAP_PTR(o->v_, o->v_s);
            sigma = sigma / mu;
            tau = tau * mu;
        }
        FN(oracle_k_energy)(masked ? mask : NULL, o->u, o->us, o->v_, o->Iu, o->Iz, o->u_last, o->a, o->b, o->c,
                            w, h, s, o->lambda, o->alpha0, o->alpha1,
                            o->mon[l][wi < ORACLE_MON_WARPS ? wi : ORACLE_MON_WARPS - 1]);
        if (masked) {
            FN(oracle_k_median5)(o->u, w, h, s, o->us);
            SWAP_PTR(o->u, o->us);
        }
        FN(oracle_k_sub)(o->u, o->u_last, w, h, s, o->du);
        FN(oracle_k_limit)(o->du, o->limitRange, w, h, s, o->du);
        FN(oracle_k_add)(o->u_last, o->du, 1, w, h, s, o->u);
        FN(oracle_k_clone)(o->u_, 1, w, h, s, o->u);
        FN(oracle_k_flowvec)(o->du, o->tv2, w, h, s, o->dwarpUV);
        FN(oracle_k_add)(o->warpUV, o->dwarpUV, 2, w, h, s, o->warpUV);
    }
}

/* Upscale u, u_ and warpUV from level l to l-1, then swap (stereotgv.cu:497-510). */
static void FN(upscale_level)(FN(oracle_t) *o, int l)
{
    int w = o->pW[l], h = o->pH[l], s = o->pS[l];
    int W = o->pW[l - 1], H = o->pH[l - 1], S = o->pS[l - 1];
    FN(oracle_k_upscale)(o->u, 1, w, h, s, W, H, S, o->fScale, o->us);
    FN(oracle_k_upscale)(o->u_, 1, w, h, s, W, H, S, o->fScale, o->u_s);
    FN(oracle_k_upscale)(o->warpUV, 2, w, h, s, W, H, S, o->fScale, o->warpUVs);
    SWAP_PTR(o->u, o->us);
    SWAP_PTR(o->u_, o->u_s);
    SWAP_PTR(o->warpUV, o->warpUVs);
}

/* Calibration warp and image pyramid up to `top` (stereotgv.cu:392-405). */
static void FN(prepare_images)(FN(oracle_t) *o, int top)
{
    FN(oracle_k_warp)(o->pI1[0], o->width, o->height, o->stride, o->cv, o->i1cal);
    SWAP_PTR(o->pI1[0], o->i1cal);
    for (int l = 1; l <= top; l++) {
        FN(oracle_k_downscale)(o->pI0[l - 1], 1, o->pW[l - 1], o->pH[l - 1], o->pS[l - 1],
                               o->pW[l], o->pH[l], o->pS[l], o->pI0[l]);
        FN(oracle_k_downscale)(o->pI1[l - 1], 1, o->pW[l - 1], o->pH[l - 1], o->pS[l - 1],
                               o->pW[l], o->pH[l], o->pS[l], o->pI1[l]);
    }
}

/* stereotgv.cu:389-515 (masked = 1) and :256-387 (masked = 0) */
int FN(oracle_solve)(FN(oracle_t) *o, int masked)
{
    FN(zero_work)(o);
    FN(prepare_images)(o, o->nLevels - 1);
    /* u, u_, warpUV were zeroed by zero_work (:395-397) */
    for (int l = o->nLevels - 1; l >= 0; l--) {
        if (l == o->nLevels - 1)
            FN(oracle_k_flowvec)(o->u, o->pTv[l], o->pW[l], o->pH[l], o->pS[l], o->warpUV);
        FN(solve_level)(o, masked, l, o->nWarpIters, o->nSolverIters);
        if (l > 0) FN(upscale_level)(o, l);
    }
    return 0;
}

/* Warm-started re-solve (no reference counterpart; the serving schedule): the previous
 * solve's u, u_ and warpUV are downscaled level by level to `start_level` with the
 * 4-tap Downscale times 1/fScale (downscale.cu:148-237) into the ping-pong partners,
 * then the solve runs from start_level to level 0 with `warps` x `iters`.  Nothing is
 * zeroed. */
int FN(oracle_solve_warm)(FN(oracle_t) *o, int masked, int start_level, int warps, int iters)
{
    if (start_level < 0 || start_level >= o->nLevels) return -1;
    FN(prepare_images)(o, start_level);
    R sc = (R)(1.0f / (float)o->fScale);
    for (int l = 1; l <= start_level; l++) {
        int W = o->pW[l - 1], H = o->pH[l - 1], S = o->pS[l - 1];
        int w = o->pW[l], h = o->pH[l], s = o->pS[l];
        FN(oracle_k_downscale)(o->u, 1, W, H, S, w, h, s, o->us);
        FN(oracle_k_downscale)(o->u_, 1, W, H, S, w, h, s, o->u_s);
        FN(oracle_k_downscale)(o->warpUV, 2, W, H, S, w, h, s, o->warpUVs);
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                size_t p = (size_t)y * s + x;
                o->us[p] *= sc;
                o->u_s[p] *= sc;
                o->warpUVs[2 * p] *= sc;
                o->warpUVs[2 * p + 1] *= sc;
            }
        SWAP_PTR(o->u, o->us);
        SWAP_PTR(o->u_, o->u_s);
        SWAP_PTR(o->warpUV, o->warpUVs);
    }
    for (int l = start_level; l >= 0; l--) {
        FN(solve_level)(o, masked, l, warps, iters);
        if (l > 0) FN(upscale_level)(o, l);
    }
    return 0;
}

/* Energy record of (level, warp) of the last solve (oracle_k_energy). */
void FN(oracle_get_energy)(const FN(oracle_t) *o, int level, int warp, double *out)
{
    if (warp >= ORACLE_MON_WARPS) warp = ORACLE_MON_WARPS - 1;
    memcpy(out, o->mon[level][warp], 5 * sizeof(double));
}

/* Outputs (stereotgv.cu:517-568), unpadded, row pitch = width. */
void FN(oracle_get_u)(const FN(oracle_t) *o, R *out)
{
    for (int y = 0; y < o->height; y++)
        memcpy(out + (size_t)y * o->width, o->u + (size_t)y * o->stride, o->width * sizeof(R));
}
void FN(oracle_get_warpuv)(const FN(oracle_t) *o, R *out)
{
    for (int y = 0; y < o->height; y++)
        memcpy(out + (size_t)y * o->width * 2, o->warpUV + (size_t)y * o->stride * 2, 2 * o->width * sizeof(R));
}
void FN(oracle_get_i1warp)(const FN(oracle_t) *o, R *out)
{
    for (int y = 0; y < o->height; y++)
        memcpy(out + (size_t)y * o->width, o->i1warp + (size_t)y * o->stride, o->width * sizeof(R));
}
void FN(oracle_stereo)(FN(oracle_t) *o, float focalx, float focaly, float cx, float cy,
                       float d1, float d2, float d3, float d4, float t1, float t2, float t3,
                       R *depth_out, R *X_out)
{
    FN(oracle_k_convert_kb)(o->warpUV, focalx, focaly, cx, cy, d1, d2, d3, d4, t1, t2, t3,
                            o->X, o->depth, o->width, o->height, o->stride);
    for (int y = 0; y < o->height; y++) {
        memcpy(depth_out + (size_t)y * o->width, o->depth + (size_t)y * o->stride, o->width * sizeof(R));
        memcpy(X_out + (size_t)y * o->width * 3, o->X + (size_t)y * o->stride * 3, 3 * o->width * sizeof(R));
    }
}

/* Named internal plane, read at `level` dims into a compact (h, w*nch) array.
 * Returns nch, or 0 for an unknown name. */
int FN(oracle_read_plane)(const FN(oracle_t) *o, const char *name, int level, R *out)
{
    const R *src = NULL;
    int nch = 1;
    if (!strcmp(name, "u")) src = o->u;
    else if (!strcmp(name, "us")) src = o->us;
    else if (!strcmp(name, "u_")) src = o->u_;
    else if (!strcmp(name, "u_s")) src = o->u_s;
    else if (!strcmp(name, "u_last")) src = o->u_last;
    else if (!strcmp(name, "a")) src = o->a;
    else if (!strcmp(name, "b")) src = o->b;
    else if (!strcmp(name, "c")) src = o->c;
    else if (!strcmp(name, "etau")) src = o->etau;
    else if (!strcmp(name, "etav1")) src = o->etav1;
    else if (!strcmp(name, "etav2")) src = o->etav2;
    else if (!strcmp(name, "Iu")) src = o->Iu;
    else if (!strcmp(name, "Iz")) src = o->Iz;
    else if (!strcmp(name, "i1warp")) src = o->i1warp;
    else if (!strcmp(name, "I0")) src = o->pI0[level];
    else if (!strcmp(name, "I1")) src = o->pI1[level];
    else if (!strcmp(name, "mask")) src = o->pMask[level];
    else if (!strcmp(name, "warpUV")) { src = o->warpUV; nch = 2; }
    else if (!strcmp(name, "tv2")) { src = o->tv2; nch = 2; }
    else if (!strcmp(name, "tv")) { src = o->pTv[level]; nch = 2; }
    else if (!strcmp(name, "p")) { src = o->p; nch = 2; }
    else if (!strcmp(name, "v_")) { src = o->v_; nch = 2; }
    else if (!strcmp(name, "Tp")) { src = o->Tp; nch = 2; }
    else if (!strcmp(name, "q")) { src = o->q; nch = 4; }
    if (!src) return 0;
    int w = o->pW[level], h = o->pH[level], s = o->pS[level];
    for (int y = 0; y < h; y++)
        memcpy(out + (size_t)y * w * nch, src + (size_t)y * s * nch, (size_t)w * nch * sizeof(R));
    return nch;
}
